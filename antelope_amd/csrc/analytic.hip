// analytic.hip - H(k + A) and J(k + A) of the analytic tight-binding classes
// (analytic.hpp) for every k point, in the layout the N-band kernels read
// from the Fourier sums: out[(slot * nk_stride + k) * N * N + ij] for the
// matrices ev.mats[slot]. The classes have no dipoles (isDipoleZero), so
// H~ = H + E.D is H.
#include "analytic.hpp"
#include "nband.hpp"

namespace sbe {
namespace {

__global__ __launch_bounds__(64) void nb_model_kernel(const NbArgs a, const NbEval ev, double2 *__restrict__ out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (k >= a.nk_stride) return;
    const int N = a.N, nn = N * N;
    const double kk[3] = {a.kpts[3 * k] + ev.A[0], a.kpts[3 * k + 1] + ev.A[1], a.kpts[3 * k + 2] + ev.A[2]};
    double2 *H = nullptr, *J[3] = {nullptr, nullptr, nullptr};
    int hcopy = -1;
    for (int m = 0; m < ev.nmat; ++m) {
        double2 *M = out + (static_cast<int64_t>(m) * a.nk_stride + k) * nn;
        for (int e = 0; e < nn; ++e) M[e] = make_double2(0.0, 0.0);
        const int t = ev.mats[m];
        if (t == kMatHtilde || t == kMatH0) {
            if (H) hcopy = m;
            else H = M;
        } else {
            J[t - kMatJx] = M;
        }
    }
    am::analytic_eval(a.model_kind, N, a.mparams, kk, H, J[0], J[1], J[2]);
    if (hcopy >= 0) {
        double2 *M = out + (static_cast<int64_t>(hcopy) * a.nk_stride + k) * nn;
        for (int e = 0; e < nn; ++e) M[e] = H[e];
    }
}

}  // namespace

hipError_t nb_launch_model(const NbArgs &a, const NbEval &ev, double2 *out, hipStream_t s)
{
    bool jx = false, jy = false;
    for (int m = 0; m < ev.nmat; ++m) {
        jx = jx || ev.mats[m] == kMatJx;
        jy = jy || ev.mats[m] == kMatJy;
    }
    if (jx != jy) return hipErrorInvalidValue;  // the classes write J_x and J_y together
    nb_model_kernel<<<dim3(static_cast<unsigned>((a.nk_stride + 63) / 64)), 64, 0, s>>>(a, ev, out);
    return hipGetLastError();
}

}  // namespace sbe
