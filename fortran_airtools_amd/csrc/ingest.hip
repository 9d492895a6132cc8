// ingest.hip -- measured transmission data to the right-hand side on the device.
//
// A FAIR tools program in file mode (main.f90:222-270) reads transmission values T (one per ray,
// p per angle, angle-major), optionally keeps every skip-th angle, and solves A x = b with
// b = -log(T) (main.f90:270).  The data go to the device once; the culled selection and the
// logarithm happen there, so the right-hand side is resident in HBM next to A.
#include "fair_internal.hpp"

namespace fair {
namespace {

constexpr int INGEST_THREADS = 256;

// b[i*p + j] = -log(T[(i*skip)*p + j]) for the nth_out kept angles (skip = 1 keeps them all).
__global__ void __launch_bounds__(INGEST_THREADS)
    k_absorbance(const double *__restrict__ T, double *__restrict__ b, int64_t p, int64_t skip, int64_t total) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
        const int64_t i = k / p, j = k - i * p;
        b[k] = -log(T[i * skip * p + j]);
    }
}

// Angles kept by the cull: 1, 1+skip, ... up to nth (none of none).
int64_t kept_angles(int64_t nth, int64_t skip) {
    if (nth <= 0) return 0;
    return skip > 0 ? (nth - 1) / skip + 1 : nth;
}

void launch_absorbance(const double *T, double *b, int64_t p, int64_t nth, int64_t skip) {
    const int64_t keep = kept_angles(nth, skip);
    const int64_t total = keep * p;
    if (total == 0) return;
    const int64_t want = (total + INGEST_THREADS - 1) / INGEST_THREADS;
    const int grid = (int)(want < 65536 ? want : 65536);
    k_absorbance<<<grid, INGEST_THREADS, 0, stream()>>>(T, b, p, skip > 0 ? skip : 1, total);
    launch_check("k_absorbance");
}

void check_sizes(int64_t p, int64_t nth, int64_t skip) {
    FAIR_CHECK(p >= 0 && nth >= 0 && skip >= 0, FAIR_ERR_ARG, "negative size");
    FAIR_CHECK(nth == 0 || p <= INT64_MAX / nth, FAIR_ERR_OVERFLOW, "p*nth overflows");
}

}  // namespace
}  // namespace fair

using namespace fair;

// main.f90:229-245 (keep angles 1, 1+skip, 1+2 skip, ... of the (p, nth) data; skip = 0 keeps
// all) followed by main.f90:270 (b = -log T).  Device pointers; b holds ceil(nth/skip)*p values.
extern "C" int fair_absorbance_dev(const double *T, int64_t p, int64_t nth, int64_t skip, double *b) {
    return guard([&] {
        check_sizes(p, nth, skip);
        FAIR_CHECK((T && b) || kept_angles(nth, skip) * p == 0, FAIR_ERR_ARG, "null argument");
        launch_absorbance(T, b, p, nth, skip);
    });
}

// The same with host arrays (returns after b is copied back).
extern "C" int fair_absorbance(const double *T, int64_t p, int64_t nth, int64_t skip, double *b) {
    return guard([&] {
        check_sizes(p, nth, skip);
        FAIR_CHECK((T && b) || kept_angles(nth, skip) * p == 0, FAIR_ERR_ARG, "null argument");
        const int64_t nb = kept_angles(nth, skip) * p;
        if (nb == 0) return;
        DBuf<double> dT((size_t)(p * nth)), db((size_t)nb);
        dT.upload(T, (size_t)(p * nth));
        launch_absorbance(dT.p, db.p, p, nth, skip);
        db.download(b, (size_t)nb);
        FAIR_HIP(hipStreamSynchronize(stream()));
    });
}
