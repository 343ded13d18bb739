#!/usr/bin/env python3
"""Extract the CRFB noise-transfer-function coefficient table into C data.

The loop-filter coefficients are data that must be identical to the
reference's for the 1-bit output to match, so they are read from the
reference's table (src/ntf.c:16-792, struct include/ntf.h:14-20) and
written as a plain C initializer list: one row per filter holding
{a[10], g[10], order, freq, name}, in the reference's table order (lookup
is "first name match with freq <= rate", src/ntf.c:814-830, so the order
matters).

Usage: gen_ntf_table.py <path to reference src/ntf.c> <output .inc>
The generated file is committed; this script is only re-run when the table
changes.
"""
import re
import sys


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _tokenize(body):
    tok = re.compile(r'\s*(\{|\}|,|"[^"]*"|[^\s{},"]+(?:\s*\*\s*[^\s{},"]+)*)')
    pos = 0
    out = []
    while pos < len(body):
        m = tok.match(body, pos)
        if not m:
            if body[pos:].strip() == "":
                break
            raise ValueError("cannot tokenize near: %r" % body[pos:pos + 40])
        out.append(m.group(1))
        pos = m.end()
    return out


def _parse(tokens, i):
    """Parse one brace group starting at tokens[i] == '{'; returns (list, next_i)."""
    assert tokens[i] == "{"
    i += 1
    items = []
    while tokens[i] != "}":
        if tokens[i] == "{":
            sub, i = _parse(tokens, i)
            items.append(sub)
        elif tokens[i] == ",":
            i += 1
            continue
        else:
            items.append(tokens[i])
            i += 1
    return items, i + 1


def _eval_int(expr):
    parts = [p.strip() for p in expr.split("*")]
    v = 1
    for p in parts:
        v *= int(p.rstrip("uU"))
    return v


def extract(path):
    text = _strip_comments(open(path).read())
    m = re.search(r"sdm_filters\[\]\s*=\s*", text)
    if not m:
        raise ValueError("table not found")
    tokens = _tokenize(text[m.end():])
    table, _ = _parse(tokens, 0)
    rows = []
    for ent in table:
        a, g, order, freq, name = ent
        a = [float(x) for x in a]
        g = [float(x) for x in g]
        rows.append((a, g, int(order), _eval_int(freq), name.strip('"')))
    return rows


def emit(rows, out_path):
    lines = [
        "/* Generated by gen_ntf_table.py from the reference CRFB table",
        " * (src/ntf.c:16-792). Data only: {a[10], g[10], order, freq, name}. */",
    ]
    for a, g, order, freq, name in rows:
        a = a + [0.0] * (10 - len(a))
        g = g + [0.0] * (10 - len(g))
        fa = ", ".join(repr(x) if x != 0 else "0.0" for x in a)
        fg = ", ".join(repr(x) if x != 0 else "0.0" for x in g)
        lines.append("{ { %s },\n  { %s },\n  %d, %du, \"%s\" }," % (fa, fg, order, freq, name))
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    rows = extract(sys.argv[1])
    emit(rows, sys.argv[2])
    print("%d filters" % len(rows))
