/* capi_driver.c -- exercises the C ABI from plain C (no Python, no torch):
 * reads a raw frame from a file, runs imop_hip_demosaic_host and writes the
 * RGB bytes to another file.  Built and run by tests/test_gpu_capi.py.
 *   capi_driver <in.raw> <out.rgb> <w> <h> <pattern> <algo> <bd> <packed> <black_level or -1> */
#include <stdio.h>
#include <stdlib.h>

#include "imop_hip.h"

int main(int argc, char** argv) {
    if (argc != 10) return 100;
    const int w = atoi(argv[3]), h = atoi(argv[4]), pat = atoi(argv[5]), algo = atoi(argv[6]), bd = atoi(argv[7]),
              packed = atoi(argv[8]);
    const float bl = (float)atof(argv[9]);
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 101;
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char* raw = (unsigned char*)malloc((size_t)n);
    if (fread(raw, 1, (size_t)n, f) != (size_t)n) return 102;
    fclose(f);
    const size_t ob = (size_t)w * h * 3 * (bd > 8 ? 2 : 1);
    unsigned char* rgb = (unsigned char*)calloc(ob, 1);
    if (imop_hip_device_count() < 1) return 103;
    const float offs[4] = {bl, bl, bl, bl};
    const int e = imop_hip_demosaic_host(raw, rgb, w, h, pat, algo, bd, packed, bl >= 0 ? offs : NULL, bl >= 0 ? 1 : 0);
    if (e != 0) return 10 + e;
    f = fopen(argv[2], "wb");
    if (!f || fwrite(rgb, 1, ob, f) != ob) return 104;
    fclose(f);
    free(raw);
    free(rgb);
    return 0;
}
