// The reference's own MSM tests, call for call, against this library: the
// flow of reference tests/test_msm.cu:132-248 (G1MSMWithGenerator),
// :366-484 (G2MSMWithGenerator) and :487-592 (G1MSMLargeN), written as a plain
// C++ program (no gtest; the reference's device.h helpers become the HIP calls
// they wrap: cuda_malloc_async -> hipMallocAsync, cuda_memset_async ->
// hipMemsetAsync, cuda_memcpy_async_to_gpu/cpu -> hipMemcpyAsync,
// cuda_synchronize_stream -> hipStreamSynchronize, cuda_drop_async ->
// hipFreeAsync).
//
// Each case takes the host generator g1_generator() / g2_generator() in
// standard form, converts a copy in place with fp_to_montgomery, fills N
// standard-form copies of G with scalars 1..N, converts them on the device
// with point_to_montgomery_batch, runs point_msm_u64_g1/g2 with scratch sized
// the reference's way ((blocks + 1) * MSM_BUCKET_COUNT projective points, far
// below what this pipeline needs, so the library must take its own), computes
// the expected N(N+1)/2 * G with point_scalar_mul_u64, normalises the MSM on
// the host with projective_to_affine_g1/g2, converts both with
// point_from_montgomery and compares them with fp_cmp / fp2_cmp. Guard
// buffers allocated around the scratch must come back untouched; whether the
// scratch itself was used is reported (the library may use it when its
// allocation is in fact large enough).
//
// usage: msm_flow [N ...]   (default: the reference's cases N = 10, 10, 100,
// then 1000 and 5000 for each curve). Prints one line per case; exit status
// 0 when every case passes, 1 otherwise, 2 on a HIP error.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "curve.h"
#include "fp.h"
#include "fp2.h"

#define HIP_OK(x)                                                          \
    do {                                                                   \
        hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) {                                            \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_)); \
            exit(2);                                                       \
        }                                                                  \
    } while (0)

static uint64_t triangular_number(uint64_t n) { return n % 2 == 0 ? (n / 2) * (n + 1) : n * ((n + 1) / 2); }

static const size_t GUARD = 1 << 16;

static void* guard_alloc(hipStream_t s) {
    void* g = nullptr;
    HIP_OK(hipMallocAsync(&g, GUARD, s));
    HIP_OK(hipMemsetAsync(g, 0x5A, GUARD, s));
    return g;
}

static bool guard_intact(void* g, hipStream_t s) {
    std::vector<unsigned char> h(GUARD);
    HIP_OK(hipMemcpyAsync(h.data(), g, GUARD, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (unsigned char c : h)
        if (c != 0x5A) return false;
    return true;
}

static void to_mont(G1Point& p) {
    fp_to_montgomery(p.x, p.x);
    fp_to_montgomery(p.y, p.y);
}
static void to_mont(G2Point& p) {
    fp_to_montgomery(p.x.c0, p.x.c0);
    fp_to_montgomery(p.x.c1, p.x.c1);
    fp_to_montgomery(p.y.c0, p.y.c0);
    fp_to_montgomery(p.y.c1, p.y.c1);
}
static const G1Point& generator(G1Point*) { return g1_generator(); }
static const G2Point& generator(G2Point*) { return g2_generator(); }
static bool is_inf(const G1Point& p) { return g1_is_infinity(p); }
static bool is_inf(const G2Point& p) { return g2_is_infinity(p); }
static void copy_coords(G1Point& d, const G1Point& s) {
    fp_copy(d.x, s.x);
    fp_copy(d.y, s.y);
    d.infinity = s.infinity;
}
static void copy_coords(G2Point& d, const G2Point& s) {
    fp2_copy(d.x, s.x);
    fp2_copy(d.y, s.y);
    d.infinity = s.infinity;
}
static void msm(hipStream_t s, G1ProjectivePoint* r, const G1Point* p, const uint64_t* k, G1ProjectivePoint* scr,
                int n) {
    point_msm_u64_g1(s, 0, r, p, k, scr, n);
}
static void msm(hipStream_t s, G2ProjectivePoint* r, const G2Point* p, const uint64_t* k, G2ProjectivePoint* scr,
                int n) {
    point_msm_u64_g2(s, 0, r, p, k, scr, n);
}
static void normalise(G1Point& a, const G1ProjectivePoint& p) { projective_to_affine_g1(a, p); }
static void normalise(G2Point& a, const G2ProjectivePoint& p) { projective_to_affine_g2(a, p); }
static bool same(const G1Point& a, const G1Point& b) {
    return a.infinity == b.infinity && (a.infinity || (fp_cmp(a.x, b.x) == 0 && fp_cmp(a.y, b.y) == 0));
}
static bool same(const G2Point& a, const G2Point& b) {
    return a.infinity == b.infinity && (a.infinity || (fp2_cmp(a.x, b.x) == 0 && fp2_cmp(a.y, b.y) == 0));
}

template <class Aff, class Proj>
static bool run_case(hipStream_t stream, uint64_t N, const char* name) {
    const Aff& G = generator((Aff*)nullptr);
    Aff G_mont = G;
    to_mont(G_mont);
    if (is_inf(G)) {
        printf("%s N=%llu: generator not set\n", name, (unsigned long long)N);
        return false;
    }
    // reference sizing (tests/test_msm.cu:146-149)
    const int threadsPerBlock = 256;
    const int num_blocks = (int)((N + threadsPerBlock - 1) / threadsPerBlock);
    const size_t scratch_size = (size_t)(num_blocks + 1) * MSM_BUCKET_COUNT * sizeof(Proj);

    void* guard_lo = guard_alloc(stream);
    Aff* d_points = nullptr;
    uint64_t* d_scalars = nullptr;
    Proj* d_result = nullptr;
    Proj* d_scratch = nullptr;
    Aff* d_expected = nullptr;
    Aff* d_G = nullptr;
    HIP_OK(hipMallocAsync((void**)&d_points, N * sizeof(Aff), stream));
    HIP_OK(hipMallocAsync((void**)&d_scalars, N * sizeof(uint64_t), stream));
    HIP_OK(hipMallocAsync((void**)&d_result, sizeof(Proj), stream));
    HIP_OK(hipMallocAsync((void**)&d_scratch, scratch_size, stream));
    void* guard_hi = guard_alloc(stream);
    HIP_OK(hipMallocAsync((void**)&d_expected, sizeof(Aff), stream));
    HIP_OK(hipMallocAsync((void**)&d_G, sizeof(Aff), stream));
    HIP_OK(hipMemsetAsync(d_result, 0, sizeof(Proj), stream));
    HIP_OK(hipMemsetAsync(d_expected, 0, sizeof(Aff), stream));
    HIP_OK(hipMemsetAsync(d_G, 0, sizeof(Aff), stream));
    HIP_OK(hipMemsetAsync(d_scratch, 0x5A, scratch_size, stream));

    std::vector<Aff> h_points(N);
    for (uint64_t i = 0; i < N; i++) copy_coords(h_points[i], G);
    std::vector<uint64_t> h_scalars(N);
    for (uint64_t i = 0; i < N; i++) h_scalars[i] = i + 1;
    HIP_OK(hipMemcpyAsync(d_points, h_points.data(), N * sizeof(Aff), hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemcpyAsync(d_scalars, h_scalars.data(), N * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    point_to_montgomery_batch<Aff>(stream, 0, d_points, (int)N);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(d_G, &G_mont, sizeof(Aff), hipMemcpyHostToDevice, stream));

    msm(stream, d_result, d_points, d_scalars, d_scratch, (int)N);
    point_scalar_mul_u64<Aff>(stream, 0, d_expected, d_G, triangular_number(N));

    Aff* d_result_affine = nullptr;
    Aff* d_expected_normal = nullptr;
    HIP_OK(hipMallocAsync((void**)&d_result_affine, sizeof(Aff), stream));
    HIP_OK(hipMallocAsync((void**)&d_expected_normal, sizeof(Aff), stream));
    HIP_OK(hipMemsetAsync(d_expected_normal, 0, sizeof(Aff), stream));
    HIP_OK(hipStreamSynchronize(stream));
    Proj h_result_proj;
    HIP_OK(hipMemcpyAsync(&h_result_proj, d_result, sizeof(Proj), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    Aff h_result_affine;
    normalise(h_result_affine, h_result_proj);
    HIP_OK(hipMemcpyAsync(d_result_affine, &h_result_affine, sizeof(Aff), hipMemcpyHostToDevice, stream));
    point_from_montgomery<Aff>(stream, 0, d_result_affine, d_result_affine);
    point_from_montgomery<Aff>(stream, 0, d_expected_normal, d_expected);
    HIP_OK(hipStreamSynchronize(stream));
    Aff msm_result, expected_result;
    HIP_OK(hipMemcpyAsync(&msm_result, d_result_affine, sizeof(Aff), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(&expected_result, d_expected_normal, sizeof(Aff), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));

    // the reference-sized scratch and the buffers around it stay untouched
    std::vector<unsigned char> scr(scratch_size);
    HIP_OK(hipMemcpyAsync(scr.data(), d_scratch, scratch_size, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    bool scratch_untouched = true;
    for (unsigned char c : scr) scratch_untouched &= c == 0x5A;
    const bool guards = guard_intact(guard_lo, stream) && guard_intact(guard_hi, stream);

    const bool ok = same(msm_result, expected_result) && !msm_result.infinity;
    printf("%s N=%llu scratch=%zu B: %s (scratch %s, guards %s)\n", name, (unsigned long long)N, scratch_size,
           ok ? "PASS" : "FAIL", scratch_untouched ? "untouched" : "WRITTEN", guards ? "intact" : "CORRUPTED");
    for (void* p : {(void*)d_points, (void*)d_scalars, (void*)d_result, (void*)d_scratch, (void*)d_expected,
                    (void*)d_G, (void*)d_result_affine, (void*)d_expected_normal, guard_lo, guard_hi})
        HIP_OK(hipFreeAsync(p, stream));
    HIP_OK(hipStreamSynchronize(stream));
    return ok && guards;
}

int main(int argc, char** argv) {
    hipStream_t stream;
    HIP_OK(hipSetDevice(0));
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    init_device_generators(stream, 0);
    HIP_OK(hipStreamSynchronize(stream));
    bool all = true;
    if (argc > 1) {
        for (int i = 1; i < argc; ++i) {
            const uint64_t n = strtoull(argv[i], nullptr, 10);
            all &= run_case<G1Point, G1ProjectivePoint>(stream, n, "G1");
            all &= run_case<G2Point, G2ProjectivePoint>(stream, n, "G2");
        }
    } else {
        all &= run_case<G1Point, G1ProjectivePoint>(stream, 10, "G1MSMWithGenerator");
        all &= run_case<G2Point, G2ProjectivePoint>(stream, 10, "G2MSMWithGenerator");
        all &= run_case<G1Point, G1ProjectivePoint>(stream, 100, "G1MSMLargeN");
        for (uint64_t n : {1000ull, 5000ull}) {
            all &= run_case<G1Point, G1ProjectivePoint>(stream, n, "G1");
            all &= run_case<G2Point, G2ProjectivePoint>(stream, n, "G2");
        }
    }
    HIP_OK(hipStreamDestroy(stream));
    printf("%s\n", all ? "ALL PASS" : "SOME FAILED");
    return all ? 0 : 1;
}
