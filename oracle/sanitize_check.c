/* Test infrastructure: runs every oracle entry point (oracle/hydro_oracle.c, both
 * precisions) on small inputs.  tests/test_oracle_sanitize.py builds it together with
 * the oracle under -fsanitize=address,undefined and runs it; it exits 0 when every
 * check passes and nothing was reported.  Usage: sanitize_check <scratch csv path> */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define real double
#define SUF(n) n##_f64
#include "sanitize_body.h"
#undef real
#undef SUF

#define real float
#define SUF(n) n##_f32
#include "sanitize_body.h"
#undef real
#undef SUF

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s <scratch csv path>\n", argv[0]);
        return 2;
    }
    const int bad = check_f64(argv[1]) + check_f32(argv[1]);
    remove(argv[1]);
    if (bad) fprintf(stderr, "sanitize_check: %d precision(s) failed\n", bad);
    return bad ? 1 : 0;
}
