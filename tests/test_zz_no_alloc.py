"""dsdhip_trellis_run_jobs_noalloc (job table in a static device array,
nothing allocated) against dsdhip_trellis_run_jobs through the C ABI: two
calls in a row on persistent contexts, the second smaller, including a job
that starts from a seeded copy; and the size limit."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pink(n, seed):
    from foo_dsd_trellis_amd import signals
    rng = np.random.default_rng(seed)
    return np.clip(0.18 * signals.pink(n, rng), -0.5, 0.5)


def _bits(d_out, kept):
    from foo_dsd_trellis_amd import trellis as T
    return [T.unpack_bits(d_out[j].cpu().numpy(), int(kept[j])) for j in range(len(kept))]


def test_noalloc_matches_run_jobs(gpu):
    from foo_dsd_trellis_amd import trellis as T
    rate, K, L, B = 11289600, 32, 64, 8
    fi = T.ntf_auto(rate)
    ref_pool = T.SdmPool(fi, B, K, L, 0.0, 4)
    na_pool = T.SdmPool(fi, B, K, L, 0.0, 4)
    seed = np.full((3, 10), np.nan)
    src = _pink(20000, 99)
    for pool in (ref_pool, na_pool):                      # context 3: the copy source, not run below
        pool.process(3, src)
    for call, n in enumerate([40000, 9000]):          # the second call is smaller
        print("case: noalloc call %d, %d samples per job" % (call, n), flush=True)
        x = np.stack([_pink(n, 10 * call + c) for c in range(3)])
        d_in = torch.from_numpy(np.ascontiguousarray(x)).cuda()
        words = n // 64 + 2
        outs = []
        for pool, noalloc in ((ref_pool, False), (na_pool, True)):
            sd = seed.copy()
            sd[2, :] = pool.best_state(3)                  # job 2: seeded copy of context 3
            d_out = torch.zeros((3, words), dtype=torch.int64, device="cuda")
            kept, slow = pool.run_jobs([0, 1, 2], [d_in[c].data_ptr() for c in range(3)], [n] * 3,
                                       [d_out[c].data_ptr() for c in range(3)], keep_from=[0, 0, 100],
                                       copy_src=[-1, -1, 3], seed=sd, noalloc=noalloc)
            torch.cuda.synchronize()
            outs.append((kept, slow, _bits(d_out, kept)))
        (k0, s0, b0), (k1, s1, b1) = outs
        assert np.array_equal(k0, k1) and np.array_equal(s0, s1), call
        for j in range(3):
            assert np.array_equal(b0[j], b1[j]), (call, j)
        for i in range(3):
            assert np.array_equal(ref_pool.best_state(i), na_pool.best_state(i)), (call, i)


def test_noalloc_rejects_more_jobs_than_its_table(gpu):
    from foo_dsd_trellis_amd import trellis as T
    from foo_dsd_trellis_amd._lib import lib
    fi = T.ntf_auto(2822400)
    pool = T.SdmPool(fi, 8, 8, 16, 0.0, 1)
    n = 4097                                                # DSDHIP_NOALLOC_MAX_JOBS + 1
    d_in = torch.zeros(64, dtype=torch.float64, device="cuda")
    d_out = torch.zeros(4, dtype=torch.int64, device="cuda")
    ci = (C.c_int * n)(*([0] * n))
    pin = (C.c_void_p * n)(*([d_in.data_ptr()] * n))
    pout = (C.c_void_p * n)(*([d_out.data_ptr()] * n))
    ni = (C.c_int64 * n)(*([64] * n))
    kept = (C.c_int64 * n)()
    rc = lib().dsdhip_trellis_run_jobs_noalloc(pool.h, n, ci, None, pin, ni, None, pout, None, kept, None, None)
    assert rc == -1
