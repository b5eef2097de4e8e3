// A C++ caller of the reference's device-pointer surface (include/curve.h),
// written the way the reference's own tests drive it (reference
// tests/test_msm.cu:170-200: device buffers, batch Montgomery conversion,
// point_msm_u64_* with caller scratch, then the host normalisation).
//
// usage: device_api g1|g2 N points.bin scalars.bin out.bin
//   points.bin  N ABI affine records in standard form (converted on the device
//               with point_to_montgomery_batch, as c_wrapper.cu:124-127 does)
//   scalars.bin N u64 scalars
//   out.bin     the affine MSM (Montgomery form), one record
// The program runs the MSM four ways (sync and async, caller scratch sized by
// point_msm_u64_scratch_bytes_* and NULL scratch), checks that all four agree
// after normalisation, and writes the result. Exit status 0 ok, 1 mismatch,
// 2 usage or IO error.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "curve.h"
#include "rust_wrapper.h"

#define HIP_OK(x)                                                        \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) {                                          \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_)); \
            exit(2);                                                     \
        }                                                                \
    } while (0)

static std::vector<unsigned char> read_file(const char* path, size_t bytes) {
    std::vector<unsigned char> b(bytes);
    FILE* f = fopen(path, "rb");
    if (!f || fread(b.data(), 1, bytes, f) != bytes) {
        fprintf(stderr, "cannot read %s\n", path);
        exit(2);
    }
    fclose(f);
    return b;
}

template <class Aff, class Proj>
static int run(int n, const char* pts_path, const char* sc_path, const char* out_path) {
    std::vector<unsigned char> pts = read_file(pts_path, (size_t)n * sizeof(Aff));
    std::vector<unsigned char> sc = read_file(sc_path, (size_t)n * 8);
    hipStream_t s;
    HIP_OK(hipStreamCreate(&s));
    Aff* d_pts;
    uint64_t* d_sc;
    Proj* d_res;
    HIP_OK(hipMalloc((void**)&d_pts, pts.size() + 1));
    HIP_OK(hipMalloc((void**)&d_sc, sc.size() + 8));
    HIP_OK(hipMalloc((void**)&d_res, 4 * sizeof(Proj)));
    HIP_OK(hipMemcpy(d_pts, pts.data(), pts.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_sc, sc.data(), sc.size(), hipMemcpyHostToDevice));
    point_to_montgomery_batch<Aff>(s, 0, d_pts, n);

    const bool g2 = sizeof(Aff) == sizeof(G2Point);
    const size_t scratch_bytes = g2 ? point_msm_u64_scratch_bytes_g2(n) : point_msm_u64_scratch_bytes_g1(n);
    void* d_scratch = nullptr;
    if (scratch_bytes) HIP_OK(hipMalloc(&d_scratch, scratch_bytes));
    for (int k = 0; k < 4; ++k) {
        Proj* scratch = (k & 1) ? nullptr : (Proj*)d_scratch;
        if (!g2) {
            if (k < 2)
                point_msm_u64_g1(s, 0, (G1ProjectivePoint*)(d_res + k), (const G1Point*)d_pts, d_sc,
                                 (G1ProjectivePoint*)scratch, n);
            else
                point_msm_u64_async_g1(s, 0, (G1ProjectivePoint*)(d_res + k), (const G1Point*)d_pts, d_sc,
                                       (G1ProjectivePoint*)scratch, n);
        } else {
            if (k < 2)
                point_msm_u64_g2(s, 0, (G2ProjectivePoint*)(d_res + k), (const G2Point*)d_pts, d_sc,
                                 (G2ProjectivePoint*)scratch, n);
            else
                point_msm_u64_async_g2(s, 0, (G2ProjectivePoint*)(d_res + k), (const G2Point*)d_pts, d_sc,
                                       (G2ProjectivePoint*)scratch, n);
        }
    }
    HIP_OK(hipStreamSynchronize(s));
    Proj h[4];
    HIP_OK(hipMemcpy(h, d_res, sizeof h, hipMemcpyDeviceToHost));
    Aff a[4];
    for (int k = 0; k < 4; ++k) {
        if (!g2)
            projective_to_affine_g1_wrapper((G1Point*)&a[k], (const G1ProjectivePoint*)&h[k]);
        else
            projective_to_affine_g2_wrapper((G2Point*)&a[k], (const G2ProjectivePoint*)&h[k]);
    }
    int rc = 0;
    for (int k = 1; k < 4; ++k)
        if (memcmp(&a[0], &a[k], offsetof(Aff, infinity) + 1) != 0) {
            fprintf(stderr, "variant %d differs from the synchronous MSM\n", k);
            rc = 1;
        }
    FILE* f = fopen(out_path, "wb");
    if (!f || fwrite(&a[0], sizeof(Aff), 1, f) != 1) return 2;
    fclose(f);
    printf("%s n=%zu scratch_bytes=%zu infinity=%d agree=%d\n", g2 ? "g2" : "g1", (size_t)n, scratch_bytes,
           (int)a[0].infinity, rc == 0);
    if (d_scratch) HIP_OK(hipFree(d_scratch));
    HIP_OK(hipFree(d_pts));
    HIP_OK(hipFree(d_sc));
    HIP_OK(hipFree(d_res));
    HIP_OK(hipStreamDestroy(s));
    return rc;
}

int main(int argc, char** argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: device_api g1|g2 N points.bin scalars.bin out.bin\n");
        return 2;
    }
    const int n = atoi(argv[2]);
    if (n < 0) return 2;
    if (strcmp(argv[1], "g2") == 0) return run<G2Point, G2ProjectivePoint>(n, argv[3], argv[4], argv[5]);
    return run<G1Point, G1ProjectivePoint>(n, argv[3], argv[4], argv[5]);
}
