/* oracle_image.c — PNG row reconstruction for the image-writer tests (TEST INFRASTRUCTURE ONLY,
 * part of liboracle.so; the product never links it).
 *
 * Undoes the five PNG filters (PNG specification, section 9: None, Sub, Up, Average, Paeth) of
 * h rows of 1 + w * bpp bytes into h rows of w * bpp bytes. Returns 0, or -1 for an unknown
 * filter type. Pinned by the test suite against stb_image's decoding of the reference's PNG
 * files (the reference's own vendored decoder). */
#include <stdint.h>
#include <stdlib.h>

static int paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

int oracle_png_unfilter(const uint8_t* in, int64_t h, int64_t w, int bpp, uint8_t* out) {
    const int64_t len = w * bpp;
    for (int64_t y = 0; y < h; ++y) {
        const uint8_t* f = in + y * (len + 1);
        uint8_t* r = out + y * len;
        const uint8_t* up = y > 0 ? r - len : NULL;
        const int type = f[0];
        for (int64_t k = 0; k < len; ++k) {
            const int x = f[1 + k];
            const int a = k >= bpp ? r[k - bpp] : 0;
            const int b = up ? up[k] : 0;
            const int c = (up && k >= bpp) ? up[k - bpp] : 0;
            int v;
            switch (type) {
                case 0: v = x; break;
                case 1: v = x + a; break;
                case 2: v = x + b; break;
                case 3: v = x + ((a + b) >> 1); break;
                case 4: v = x + paeth(a, b, c); break;
                default: return -1;
            }
            r[k] = (uint8_t)v;
        }
    }
    return 0;
}
