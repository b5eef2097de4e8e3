/*
 * What the reference's Rust layer does around the C ABI, in C11 (which also
 * shows that rust_wrapper.h is a plain C header):
 *   G1Projective::msm / G2Projective::msm  (reference rust-api/src/types.rs:345-376, 492-532)
 *   .to_affine()                           (types.rs:331-341, 472-490)
 *   g1/g2_affine_from_montgomery           (rust-api/src/conversions.rs:83-106)
 * through the extern block of rust-api/src/ffi.rs:64-107.
 *
 * usage: rust_flow g1|g2 N points.bin scalars.bin out.bin
 *   points.bin   N ABI affine records as a Rust caller holds them (G1 in
 *                standard form, G2 in Montgomery form: c_wrapper.cu:124-127, 223-224)
 *   scalars.bin  N u64 scalars
 *   out.bin      the affine result in normal (non-Montgomery) form, one record
 * Exit status: 0 ok, 2 usage/IO error, 3 the wrapper returned an error
 * (its code is printed, as the Rust layer's Err(String) would carry it).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rust_wrapper.h"

static void* read_file(const char* path, size_t bytes) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    void* buf = malloc(bytes ? bytes : 1);
    size_t got = bytes ? fread(buf, 1, bytes, f) : 0;
    fclose(f);
    if (got != bytes) {
        free(buf);
        return NULL;
    }
    return buf;
}

static int write_file(const char* path, const void* p, size_t bytes) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    size_t put = fwrite(p, 1, bytes, f);
    fclose(f);
    return put == bytes ? 0 : -1;
}

int main(int argc, char** argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: rust_flow g1|g2 N points.bin scalars.bin out.bin\n");
        return 2;
    }
    const int g2 = strcmp(argv[1], "g2") == 0;
    const int n = atoi(argv[2]);
    const size_t rec = g2 ? sizeof(G2Point) : sizeof(G1Point);
    void* pts = read_file(argv[3], (size_t)n * rec);
    uint64_t* sc = (uint64_t*)read_file(argv[4], (size_t)n * sizeof(uint64_t));
    if (n < 0 || !pts || !sc) {
        fprintf(stderr, "rust_flow: cannot read inputs\n");
        return 2;
    }
    if (!g2) {
        G1ProjectivePoint acc;
        G1Point aff, out;
        if (n == 0) {
            g1_projective_point_at_infinity_wrapper(&acc); /* types.rs:351-353: empty -> infinity */
        } else {
            int rc = g1_msm_wrapper(&acc, (const G1Point*)pts, sc, n, 0);
            if (rc != 0) {
                printf("MSM computation failed with error code: %d\n", rc);
                return 3;
            }
        }
        projective_to_affine_g1_wrapper(&aff, &acc);
        if (g1_is_infinity_wrapper(&aff))
            g1_point_at_infinity_wrapper(&out);
        else
            g1_from_montgomery_wrapper(&out, &aff);
        if (write_file(argv[5], &out, sizeof out)) return 2;
        printf("g1 n=%d infinity=%d\n", n, (int)out.infinity);
    } else {
        G2ProjectivePoint acc;
        G2Point aff, out;
        if (n == 0) {
            g2_projective_point_at_infinity_wrapper(&acc);
        } else {
            int rc = g2_msm_wrapper(&acc, (const G2Point*)pts, sc, n, 0);
            if (rc != 0) {
                printf("MSM computation failed with error code: %d\n", rc);
                return 3;
            }
        }
        projective_to_affine_g2_wrapper(&aff, &acc);
        if (g2_is_infinity_wrapper(&aff))
            g2_point_at_infinity_wrapper(&out);
        else
            g2_from_montgomery_wrapper(&out, &aff);
        if (write_file(argv[5], &out, sizeof out)) return 2;
        printf("g2 n=%d infinity=%d\n", n, (int)out.infinity);
    }
    free(pts);
    free(sc);
    return 0;
}
