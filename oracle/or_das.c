/* placeholder, filled below */
#include "dsd_oracle.h"
