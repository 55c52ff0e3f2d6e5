// haldane_adaptive.hip - adaptive Cash-Karp 5(4) propagation of the 2-band
// (Haldane) SBE, single solves and batches of solves.
//
// The reference integrates with a fixed step (SBEs.h:596-624); this stepper
// is an addition with no counterpart there. It keeps the reference's model
// (Materials/Haldane.h:86-170), right-hand side (SBEs.h:762-861) and time
// samples (antelope.cpp:297-338: the currents at t_n = atmin + n dt, from the
// state at t_n), and replaces the step sequence:
//
//  * Step control per wave. The 64 k points of one wave share a step size,
//    chosen by the standard controller h <- h min(5, max(0.2, 0.9 err^-1/5))
//    from err = max over the wave's k points and state components of
//    |y5 - y4| / (atol + rtol max(|y|, |y5|)); a step with err > 1 is
//    rejected and retried. The k points are independent ODEs that share only
//    the laser, so this meets the tolerance for every k point of the solve,
//    as an instance-wide controller would, with no synchronisation between
//    waves. The error estimate and the controller run in FP64 whatever the
//    state precision.
//  * Laser at the stage times, in FP64, interpolated between the samples of
//    a host-built table of A and dA/dt and turned from the sample's phase
//    row (laser_rows), for all six stage times of a step at once across the
//    lanes of the wave.
//  * Samples. The k-summed observables (inter/intra currents, populations,
//    rho sums) are needed every dt, while a step covers several dt. At every
//    step end the kernel forms the observables of each k point at the step
//    end (the phases and d of the stage at c = 1, which is evaluated at the
//    step end itself and also gives the next step's first stage), reduces
//    them over the wave, and keeps the last six step ends in a ring; a
//    sample is the Lagrange interpolant through the ring's nodes (centred:
//    samples up to node c - 2 are written when node c arrives; one-sided at
//    the window ends). A sample then costs a few FMAs per wave instead of a
//    per-k evaluation. On the batched workload's worst solve this is as
//    accurate as a cubic Hermite interpolant with exact time derivatives at
//    two step ends, at a fraction of the cost (no derivatives).
//  * FP32 state. Near the ground state s = -d/|d| + delta with delta small,
//    so 2 d x s is a difference of O(|d|) products that FP32 resolves only
//    to ~1e-7 |d| absolute. The FP32 kernel therefore propagates delta =
//    s + dh with the reference vector dh = d / sqrt(|d|^2 + eps^2), parallel
//    to d(k + A(t)): d x dh = 0 and the dephasing acts on the part normal to
//    d, so d(delta)/dt = f(delta) + d(dh)/dt with f the right-hand side of
//    s itself (2 d x delta less the dephasing of the part of delta normal to
//    d): the same ODE with no cancelling terms (d(dh)/dt from the k gradient
//    of d along dA/dt). eps (2 % of a bound on |d|) keeps d(dh)/dt bounded
//    where the gap closes; for |d| << eps the form reverts to s itself, whose
//    FP32 error is small there (it scales with |d|). s <-> delta are
//    converted in FP64 at the window ends, and the observables are formed in
//    FP64 from s = delta - dh with FP64 k phases. The FP64 kernel propagates
//    s as the oracle restates it.
//  * A launch covers the samples [out0, out1) and ends with the state exactly
//    at t_out1 (the last step is shortened to land there), so consecutive
//    launches continue seamlessly and the partial-sum buffer stays bounded.
#include "haldane_dev.hpp"

namespace sbe {
namespace {

using namespace dev;

// Cash & Karp, ACM TOMS 16 (1990) 201, Table 1 (the host Tableau holds the
// same pair).
struct CashKarp {
    static constexpr double c[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
    static constexpr double a[6][5] = {
        {0, 0, 0, 0, 0},
        {0.2, 0, 0, 0, 0},
        {3. / 40., 9. / 40., 0, 0, 0},
        {0.3, -0.9, 1.2, 0, 0},
        {-11. / 54., 2.5, -70. / 27., 35. / 27., 0},
        {1631. / 55296., 175. / 512., 575. / 13824., 44275. / 110592., 253. / 4096.}};
    static constexpr double b[6] = {37. / 378., 0, 250. / 621., 125. / 594., 0, 512. / 1771.};
    // b - bhat (5th- minus 4th-order weights)
    static constexpr double e[6] = {37. / 378. - 2825. / 27648., 0, 250. / 621. - 18575. / 48384.,
                                    125. / 594. - 13525. / 55296., -277. / 14336., 512. / 1771. - 0.25};
};

__device__ __forceinline__ double uniform(double x)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// x of lane l
__device__ __forceinline__ double lane_value(double x, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}

// 1/x and 1/sqrt(x) for x > 0: in FP64 the hardware estimate and two Newton
// steps (full precision up to the last bit or two), in FP32 the IEEE forms
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ float recip(float x) { return 1.0f / x; }
__device__ __forceinline__ double rsqrt_fast(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    return y * fma(-0.5 * x * y, y, 1.5);
}
__device__ __forceinline__ float rsqrt_fast(float x) { return 1.0f / sqrtf(x); }

// max that propagates NaN (fmax would drop it and accept a non-finite step)
__device__ __forceinline__ double nanmax(double a, double b) { return (a != a || a > b) ? a : b; }

// Maximum over the wave of values that are >= 0 or +inf (never NaN), in every
// lane and in scalar registers: wave_max with the hardware maximum.
__device__ __forceinline__ double wave_max_nonneg(double x);

// Maximum over the wave (NaN if any lane holds one), in every lane and in
// scalar registers: DPP within each row of 16 lanes, then the row swaps, so
// no step waits on an LDS permute. Requires all 64 lanes.
__device__ __forceinline__ double wave_max(double x)
{
    x = nanmax(x, dpp_f64<0xB1>(x));   // quad_perm [1,0,3,2]
    x = nanmax(x, dpp_f64<0x4E>(x));   // quad_perm [2,3,0,1]
    x = nanmax(x, dpp_f64<0x141>(x));  // row_half_mirror: the other quad of the half row
    x = nanmax(x, dpp_f64<0x140>(x));  // row_mirror: the other half row
    double a = x, b = x;
    swap16(a, b);  // a: rows 0, 0, 2, 2; b: rows 1, 1, 3, 3
    x = nanmax(a, b);
    a = x;
    b = x;
    swap32(a, b);  // a: rows 0, 1, 0, 1; b: rows 2, 3, 2, 3
    return uniform(nanmax(a, b));
}

__device__ __forceinline__ double wave_max_nonneg(double x)
{
    x = fmax(x, dpp_f64<0xB1>(x));   // quad_perm [1,0,3,2]
    x = fmax(x, dpp_f64<0x4E>(x));   // quad_perm [2,3,0,1]
    x = fmax(x, dpp_f64<0x141>(x));  // row_half_mirror
    x = fmax(x, dpp_f64<0x140>(x));  // row_mirror
    double a = x, b = x;
    swap16(a, b);
    x = fmax(a, b);
    a = x;
    b = x;
    swap32(a, b);
    return uniform(fmax(a, b));
}

constexpr int kRow = 14;  // 12 phase values, then dA/dt (x, y)
// c of the Cash-Karp stages 2..6 by laser row (row 3 is c = 1 and row 5 the
// step end: both are taken at the step end itself)
__constant__ double kRowC[6] = {0.2, 0.3, 0.6, 1.0, 0.875, 1.0};
static_assert(CashKarp::c[1] == 0.2 && CashKarp::c[2] == 0.3 && CashKarp::c[3] == 0.6 && CashKarp::c[4] == 1.0 &&
                  CashKarp::c[5] == 0.875,
              "laser rows follow the Cash-Karp stage times");

// sin and cos of a small angle (|x| <= 0.25) by their Taylor series to x^13
// and x^12 (Horner in x^2 with the inverse factorials as constants); the
// first omitted terms are below 1e-18 relative.
// A double constant formed in scalar registers where it is used (two
// s_mov_b32 per use). Loop-invariant literals would otherwise be kept in
// vector registers for the whole loop, which pushes the stepper into spills.
template <unsigned long long B>
__device__ __forceinline__ double sconst()
{
    int lo, hi;
    asm volatile("s_mov_b32 %0, %1" : "=s"(lo) : "i"(static_cast<int>(B & 0xffffffffull)));
    asm volatile("s_mov_b32 %0, %1" : "=s"(hi) : "i"(static_cast<int>(B >> 32)));
    return __hiloint2double(hi, lo);
}
#define SBE_SCONST(v) sconst<__builtin_bit_cast(unsigned long long, static_cast<double>(v))>()

// sin and cos of |x| <= 0.25 by their Taylor series (to x^13 and x^12: the
// first omitted terms are below 1e-19 relative). The series in x^2 are summed
// from the smallest term up, with the powers of x^2 formed once for both and
// every coefficient a multiplicand (a scalar-register operand, no copies).
__device__ __forceinline__ void sincos_small(double x, double &sn, double &cs)
{
    const double x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x4 * x4, x10 = x8 * x2, x12 = x8 * x4;
    double ps = SBE_SCONST(1.0 / 6227020800.0) * x12;  // 1/13!
    ps = fma(SBE_SCONST(-1.0 / 39916800.0), x10, ps);
    ps = fma(SBE_SCONST(1.0 / 362880.0), x8, ps);
    ps = fma(SBE_SCONST(-1.0 / 5040.0), x6, ps);
    ps = fma(SBE_SCONST(1.0 / 120.0), x4, ps);
    ps = fma(SBE_SCONST(-1.0 / 6.0), x2, ps);  // sin x / x - 1
    sn = fma(ps, x, x);
    double pc = SBE_SCONST(1.0 / 479001600.0) * x12;  // 1/12!
    pc = fma(SBE_SCONST(-1.0 / 3628800.0), x10, pc);
    pc = fma(SBE_SCONST(1.0 / 40320.0), x8, pc);
    pc = fma(SBE_SCONST(-1.0 / 720.0), x6, pc);
    pc = fma(SBE_SCONST(1.0 / 24.0), x4, pc);
    cs = fma(-0.5, x2, pc) + 1.0;  // cos x
}

// sin and cos of a larger angle, out of line: only pulses that move A.v by
// more than 0.25 rad within one sample interval take it, and its argument
// reduction would otherwise hold registers in every caller
__device__ __attribute__((noinline)) void sincos_large(double x, double *sn, double *cs) { sincos(x, sn, cs); }

// x^(-1/5) for the step-size controller: FP32 hardware log2/exp2 for the
// first estimate, then two Newton steps in FP64 (y <- y (6 - 5 x y^5) / ... :
// y += y (1 - x y^5) / 5), within an ulp or two of pow(x, -0.2)
__device__ __forceinline__ double inv_fifth_root(double x)
{
    const float lx = __builtin_amdgcn_logf(static_cast<float>(x));
    double y = static_cast<double>(__builtin_amdgcn_exp2f(-0.2f * lx));
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double y2 = y * y;
        y = fma(y * (1.0 - x * y2 * y2 * y), 0.2, y);
    }
    return y;
}

// What one lane needs for the laser rows, read once per launch: the lane's
// time (stage fraction, or the step end) and lattice vector, its column of
// the phase table, and the table constants. The step loop then reads only
// the two table rows its times fall between, with one memory round trip.
struct LaserLane {
    const double *a_tab;      // [nt][4]: A - A(t_min) (x, y) and dA/dt (x, y) at t_n
    const double *phase_col;  // phase_tab + 2 j: scaled e^{iA(t_n).v_j} at row n x 12
    double atmin, h, inv_h;   // sample grid
    double c;                 // stage fraction of time s (kRowC)
    double vx, vy;            // v_j: a_j (j < 3) or b_{j-3}
    int nt2;                  // last interval index (nt - 2)
    int s, j;                 // time s, phase j (lanes < 36)
    bool active, end;         // lane < 36; time s is the step end itself (s = 3, 5)
};

__device__ __forceinline__ LaserLane laser_lane(const AdaptiveArgs &A)
{
    LaserLane L;
    const int lane = __lane_id();
    L.s = lane / 6;
    L.j = lane - 6 * L.s;
    L.active = lane < 36;
    // s = 5, and the stage at c = 1 (s = 3): the step end itself
    L.end = L.s == 3 || L.s == 5;
    L.c = kRowC[L.active ? L.s : 0];
    const HaldaneArgs &a = A.h;
    const double *vrow = L.j < 3 ? a.va[L.j] : a.vb[L.j - 3];
    L.vx = vrow[0];
    L.vy = vrow[1];
    L.a_tab = A.a_tab;
    L.phase_col = a.phase_tab + 2 * L.j;
    L.atmin = A.atmin;
    L.h = A.dt_out;
    L.inv_h = A.inv_dt_out;
    L.nt2 = static_cast<int>(A.nt) - 2;  // the host keeps nt < 2^31, so 32-bit indices
    return L;
}

// Laser rows of six times (the stages 2..6 of a Cash-Karp step and its end
// t + h, or six copies of one time) in the layout of the sample table:
// t1 e^{iA.a_j} (j = 0..2) then e^{iA.b_j}, as (re, im) pairs, followed by
// dA/dt. Lane 6s + j (36 lanes) makes the phase j of time s: A between the
// samples t_i <= tau <= t_i+1 is the cubic Hermite interpolant of A and dA/dt
// at the two samples (error <= dt^4 max|d4A/dt4| / 384: 4e-13 relative for
// a 0.014 au carrier at dt = 0.25), dA/dt its derivative, and the phase is
// the sample row's e^{iA(t_i).v} turned by the small angle (A(tau) - A(t_i)).v.
// The rows go to this wave's LDS slice, from which every lane reads them as
// broadcasts.
template <typename R>
__device__ __forceinline__ void laser_rows(const LaserLane &L, double t, double hh, double tn, R *rows /*[6][kRow]*/,
                                           double *row64 = nullptr /*[6][kRow]: FP64 copies of the rows*/)
{
    const int s = L.s, j = L.j;
    if (L.active) {
        const double tau = L.end ? tn : fma(L.c, hh, t);
        // interval of the sample table
        const double h = L.h, inv_h = L.inv_h;
        const double x = (tau - L.atmin) * inv_h;
        int i = static_cast<int>(floor(x));
        i = i > L.nt2 ? L.nt2 : i;
        i = i < 0 ? 0 : i;  // the table holds at least two samples
        const double *at = L.a_tab + 4 * i;  // A_x, A_y, dA_x/dt, dA_y/dt at t_i, then at t_i+1
        double T[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) T[q] = at[q];
        const double *ph = L.phase_col + 12 * i;  // scaled e^{iA(t_i).v_j}, read with the A rows
        const double ph0 = ph[0], ph1 = ph[1];
        const double th = x - static_cast<double>(i), th2 = th * th, th3 = th2 * th;
        const double h01 = 3.0 * th2 - 2.0 * th3, h10 = (th3 - 2.0 * th2 + th) * h, h11 = (th3 - th2) * h;
        const double g01 = (6.0 * th - 6.0 * th2) * inv_h, g10 = 3.0 * th2 - 4.0 * th + 1.0, g11 = 3.0 * th2 - 2.0 * th;
        const double dAx = T[4] - T[0], dAy = T[5] - T[1];
        const double ux = h01 * dAx + h10 * T[2] + h11 * T[6];  // A(tau) - A(t_i)
        const double uy = h01 * dAy + h10 * T[3] + h11 * T[7];
        const double gx = g01 * dAx + g10 * T[2] + g11 * T[6];  // dA/dt at tau
        const double gy = g01 * dAy + g10 * T[3] + g11 * T[7];
        const double del = ux * L.vx + uy * L.vy;
        double sn, cs;
        if (fabs(del) <= 0.25) sincos_small(del, sn, cs);
        else sincos_large(del, &sn, &cs);
        const double re = ph0 * cs - ph1 * sn, im = ph1 * cs + ph0 * sn;
        rows[s * kRow + 2 * j] = static_cast<R>(re);
        rows[s * kRow + 2 * j + 1] = static_cast<R>(im);
        if (j == 0) {
            rows[s * kRow + 12] = static_cast<R>(gx);
            rows[s * kRow + 13] = static_cast<R>(gy);
        }
        if (row64) {
            row64[s * kRow + 2 * j] = re;
            row64[s * kRow + 2 * j + 1] = im;
            if (j == 0) {
                row64[s * kRow + 12] = gx;
                row64[s * kRow + 13] = gy;
            }
        }
    }
    // the writes of other lanes of this wave are read next: LDS operations of
    // one wave complete in order; keep the compiler from moving reads above
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Weighted observables of one k point at a step end and their time
// derivatives: o = w x (inter_x, inter_y, intra_x, intra_y, s_x, s_y, s_z,
// s.dhat) exactly as the fixed-step sample (antelope.cpp:307-333,
// SBEs.h:694-760 in Bloch form), and do/dt along the trajectory, with
// ds/dt = f and d/dt g(k + A(t)) = grad_k g . dA/dt (the k Hessian of d enters
// through the currents). Stage-precision arithmetic; FP64 accumulation.
template <typename R>
__device__ __forceinline__ void observe(const KPoint<R> &p, R nn, const Vec3<R> &s, const Vec3<R> &f, const R *row,
                                        const HaldaneArgs &a, R ii, double o[8], double od[8])
{
    R ar[3], ai[3], br[3], bi[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        cmul(p.er[j], p.ei[j], row[2 * j], row[2 * j + 1], ar[j], ai[j]);  // t1 e^{i(k+A).a_j}
        cmul(p.er[j + 3], p.ei[j + 3], row[6 + 2 * j], row[7 + 2 * j], br[j], bi[j]);
    }
    const R c3 = static_cast<R>(a.c3), c0 = static_cast<R>(a.c0);
    const R ux = row[12], uy = row[13];
    const Vec3<R> d{ar[0] + ar[1] + ar[2], ai[0] + ai[1] + ai[2], static_cast<R>(a.M0) - c3 * (bi[0] + bi[1] + bi[2])};
    // d_a of (d_x, d_y, d_z, d0) for a = x, y; dd/dt; d/dt of the d_a's
    R xx = 0, xy = 0, yx = 0, yy = 0, zx = 0, zy = 0, ox = 0, oy = 0;
    R tdx = 0, tdy = 0, tdz = 0;
    R gxx = 0, gxy = 0, gyx = 0, gyy = 0, gzx = 0, gzy = 0, gox = 0, goy = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const R ax = static_cast<R>(a.va[j][0]), ay = static_cast<R>(a.va[j][1]);
        const R bx = static_cast<R>(a.vb[j][0]), by = static_cast<R>(a.vb[j][1]);
        const R al = ax * ux + ay * uy, be = bx * ux + by * uy;  // a_j . dA/dt, b_j . dA/dt
        xx -= ax * ai[j];
        xy -= ay * ai[j];
        yx += ax * ar[j];
        yy += ay * ar[j];
        zx += bx * br[j];
        zy += by * br[j];
        ox += bx * bi[j];
        oy += by * bi[j];
        tdx -= al * ai[j];
        tdy += al * ar[j];
        tdz += be * br[j];
        gxx -= ax * al * ar[j];
        gxy -= ay * al * ar[j];
        gyx -= ax * al * ai[j];
        gyy -= ay * al * ai[j];
        gzx += bx * be * bi[j];
        gzy += by * be * bi[j];
        gox += bx * be * br[j];
        goy += by * be * br[j];
    }
    zx *= -c3;
    zy *= -c3;
    ox *= -c0;
    oy *= -c0;
    tdz *= -c3;
    gzx *= c3;
    gzy *= c3;
    gox *= -c0;
    goy *= -c0;
    const Vec3<R> td{tdx, tdy, tdz};
    const R dd = d.x * d.x + d.y * d.y + d.z * d.z;
    const R inv = dd > R(0) ? recip(dd) : R(0);
    const R ddt = d.x * td.x + d.y * td.y + d.z * td.z;  // (d.dd/dt) = (d|d|^2/dt)/2
    const R inv_t = R(-2) * inv * inv * ddt;
    const R sd = s.x * d.x + s.y * d.y + s.z * d.z;
    const R sd_t = f.x * d.x + f.y * d.y + f.z * d.z + s.x * td.x + s.y * td.y + s.z * td.z;
    // per axis: t = s.d_a d, q = d.d_a d, p = sd q / |d|^2, and their rates
    auto axis = [&](R gx, R gy, R gz, R hx, R hy, R hz, R &t, R &t_t, R &pp, R &pp_t) {
        t = s.x * gx + s.y * gy + s.z * gz;
        t_t = f.x * gx + f.y * gy + f.z * gz + s.x * hx + s.y * hy + s.z * hz;
        const R q = d.x * gx + d.y * gy + d.z * gz;
        const R q_t = td.x * gx + td.y * gy + td.z * gz + d.x * hx + d.y * hy + d.z * hz;
        pp = sd * q * inv;
        pp_t = (sd_t * q + sd * q_t) * inv + sd * q * inv_t;
    };
    R tx, tx_t, px, px_t, ty, ty_t, py, py_t;
    axis(xx, yx, zx, gxx, gyx, gzx, tx, tx_t, px, px_t);
    axis(xy, yy, zy, gxy, gyy, gzy, ty, ty_t, py, py_t);
    const R ni = R(1) - ii;
    const double w = p.w;
    o[0] -= static_cast<double>(tx - ii * px + ni * nn * ox) * w;
    o[1] -= static_cast<double>(ty - ii * py + ni * nn * oy) * w;
    o[2] -= static_cast<double>(ii * (nn * ox + px)) * w;
    o[3] -= static_cast<double>(ii * (nn * oy + py)) * w;
    o[4] += static_cast<double>(s.x) * w;
    o[5] += static_cast<double>(s.y) * w;
    o[6] += static_cast<double>(s.z) * w;
    od[0] -= static_cast<double>(tx_t - ii * px_t + ni * nn * gox) * w;
    od[1] -= static_cast<double>(ty_t - ii * py_t + ni * nn * goy) * w;
    od[2] -= static_cast<double>(ii * (nn * gox + px_t)) * w;
    od[3] -= static_cast<double>(ii * (nn * goy + py_t)) * w;
    od[4] += static_cast<double>(f.x) * w;
    od[5] += static_cast<double>(f.y) * w;
    od[6] += static_cast<double>(f.z) * w;
    const bool deg = !(dd > R(0));  // degenerate point: eigenbasis = orbital basis
    const R rinv = deg ? R(0) : rsqrt_fast(dd);
    o[7] += static_cast<double>(deg ? -s.z : sd * rinv) * w;
    od[7] += static_cast<double>(deg ? -f.z : sd_t * rinv - sd * rinv * inv * ddt) * w;
}

// d(k + A) from the k phases and a stage row (same arithmetic as the fixed-step kernel)
template <typename R>
__device__ __forceinline__ Vec3<R> stage_d(const KPoint<R> &p, const R *row, R c3, R M0)
{
    R dx = 0, dy = 0, sb = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        dx += p.er[j] * row[2 * j] - p.ei[j] * row[2 * j + 1];
        dy += p.er[j] * row[2 * j + 1] + p.ei[j] * row[2 * j];
        sb += p.er[j + 3] * row[7 + 2 * j] + p.ei[j + 3] * row[6 + 2 * j];
    }
    return {dx, dy, M0 - c3 * sb};
}

// d(k + A) and dd/dt = grad_k d . dA/dt from k phases (er, ei) and a row
// (with dA/dt in row[12..13]); any arithmetic type T.
template <typename T, typename P>
__device__ __forceinline__ void d_and_rate(const P *er, const P *ei, const T *row, const HaldaneArgs &a, Vec3<T> &d,
                                           Vec3<T> &td)
{
    const T c3 = static_cast<T>(a.c3), ux = row[12], uy = row[13];
    T dx = 0, dy = 0, sb = 0, tx = 0, ty = 0, tz = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T ar, ai, br, bi;
        cmul(static_cast<T>(er[j]), static_cast<T>(ei[j]), row[2 * j], row[2 * j + 1], ar, ai);
        cmul(static_cast<T>(er[j + 3]), static_cast<T>(ei[j + 3]), row[6 + 2 * j], row[7 + 2 * j], br, bi);
        const T al = static_cast<T>(a.va[j][0]) * ux + static_cast<T>(a.va[j][1]) * uy;
        const T be = static_cast<T>(a.vb[j][0]) * ux + static_cast<T>(a.vb[j][1]) * uy;
        dx += ar;
        dy += ai;
        sb += bi;
        tx -= al * ai;
        ty += al * ar;
        tz += be * br;
    }
    d = {dx, dy, static_cast<T>(a.M0) - c3 * sb};
    td = {tx, ty, -c3 * tz};
}

// Reference vector dh = d / sqrt(|d|^2 + eps^2) (the unit vector d/|d| where
// |d| >> eps, going smoothly to 0 at a gap closing) and its rate
// (dd/dt - dh (dh . dd/dt)) / sqrt(|d|^2 + eps^2).
template <typename T>
__device__ __forceinline__ void unit_and_rate(const Vec3<T> &d, const Vec3<T> &td, T eps2, Vec3<T> &dh, Vec3<T> &dht)
{
    const T r = rsqrt_fast(d.x * d.x + d.y * d.y + d.z * d.z + eps2);
    dh = {d.x * r, d.y * r, d.z * r};
    const T p = dh.x * td.x + dh.y * td.y + dh.z * td.z;
    dht = {(td.x - dh.x * p) * r, (td.y - dh.y * p) * r, (td.z - dh.z * p) * r};
}

// f/2 = d x y - (g2/2) y + ((g2 - g1)/2) (y.d / |d|^2) d  (T1 off: g1 = 0)
template <typename R, bool DEPH>
__device__ __forceinline__ Vec3<R> half_f(const Vec3<R> &d, const Vec3<R> &y, R gh2, R gh12)
{
    Vec3<R> f;
    f.x = d.y * y.z - d.z * y.y;
    f.y = d.z * y.x - d.x * y.z;
    f.z = d.x * y.y - d.y * y.x;
    if (DEPH) {
        const R dd = d.x * d.x + d.y * d.y + d.z * d.z;
        const R inv = dd > R(0) ? recip(dd) : R(0);
        const R c = -gh12 * (y.x * d.x + y.y * d.y + y.z * d.z) * inv;
        f.x += c * d.x - gh2 * y.x;
        f.y += c * d.y - gh2 * y.y;
        f.z += c * d.z - gh2 * y.z;
    }
    return f;
}

template <typename R>
__device__ __forceinline__ void kahan(R &sum, R &c, R inc)
{
    const R y = inc - c;
    const R t = sum + y;
    c = (t - sum) - y;
    sum = t;
}

// Samples from the last six step ends of a wave: the interpolating
// polynomial through them (fewer than six right after a window start) in
// Newton form on the newest nodes,
//   p(tau) = D_0 + (tau - t_n)(D_1 + (tau - t_n-1)(D_2 + ... (tau - t_n-4) D_5)),
// with D_k = f[t_n, ..., t_n-k] the backward divided differences. A new node
// updates them in place, D'_0 = f_new, D'_k = (D'_k-1 - D_k-1) / (t_new - t_n+1-k),
// and the oldest node drops out of the stencil by itself. The node times are
// uniform (in the wave's LDS slice); each lane keeps the differences of its
// wave-summed observable (value lane >> 3).
struct SampleRing {
    double D[6];  // divided differences of this lane's value, newest node first
    double *t;    // [6] node times, newest first (LDS)
    int cnt;      // nodes pushed in this window

    __device__ __forceinline__ void push(double tn, double val)
    {
        // 1/(tn - t[k]) for k < 5 (wave-uniform): lane k forms the k-th, the
        // other lanes read it from there. t[k] is node n - k (before the shift
        // below); nodes not yet pushed give 0 and an unused factor.
        const int lane = __lane_id();
        const double rl = recip(tn - t[lane < 5 ? lane : 0]);
        double prev = D[0];
        D[0] = val;
#pragma unroll
        for (int k = 1; k < 6; ++k) {
            const double old = D[k];
            D[k] = k <= cnt ? (D[k - 1] - prev) * lane_value(rl, k - 1) : 0.0;
            prev = old;
        }
        double tt[6];
        tt[0] = tn;
#pragma unroll
        for (int i = 1; i < 6; ++i) tt[i] = t[i - 1];
        ++cnt;
        // every lane holds the same times: the stores agree; reads come after the wave barrier
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < 6; ++i) t[i] = tt[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // p(tau), with the node times tk[6] read from t beforehand
    __device__ __forceinline__ double at(double tau, const double tk[6]) const
    {
        double acc = D[5];
#pragma unroll
        for (int k = 4; k >= 0; --k) acc = fma(acc, tau - tk[k], D[k]);
        return acc;
    }

    // Writes the samples mi, mi + 1, ... (below mend) whose times
    // atmin + n dt_out do not exceed t_hi, and returns the next one. The eight
    // lanes of a group hold the same value's ring, so lane l evaluates sample
    // mi + (l & 7) of value l >> 3: eight samples per pass, each with the
    // arithmetic of a single evaluation. part points at this lane's value
    // column of sample row out0.
    __device__ __forceinline__ int emit(int mi, int mend, double atmin, double dt_out, double t_hi, double *part,
                                        int64_t out0, int64_t row_stride) const
    {
        double tk[6];  // the node times, from LDS once per call
#pragma unroll
        for (int k = 0; k < 6; ++k) tk[k] = t[k];
        const int j = __lane_id() & 7;
        while (mi < mend) {
            const bool in = j < mend - mi;
            const int mj = in ? mi + j : mi;
            const double tm = atmin + dt_out * static_cast<double>(mj);
            const bool ok = in && !(tm > t_hi);
            // the times grow with j, so the lanes that write are a prefix of each group
            const int n = __popcll(__ballot(ok) & 0xffull);
            if (ok) part[(static_cast<int64_t>(mj) - out0) * row_stride] = at(tm, tk);
            mi += n;
            if (n < 8) break;
        }
        return mi;
    }
};

// One wave owns the 64 * KPT consecutive k points [64 KPT w, 64 KPT (w+1))
// of its solve (lane l: points 64 q + l, q < KPT); the wave is the unit of
// step control. KPT = 2 gives each lane two independent dependency chains
// and halves the per-wave work (laser, controller, reductions) per k point.
template <typename R, bool DEPH, int KPT>
__device__ void adaptive_body(const AdaptiveArgs &A, int64_t blk, const int64_t out0, const int64_t out1)
{
    using CK = CashKarp;
    constexpr bool kComp = StateT<R>::kComp != 0;
    const HaldaneArgs &a = A.h;
    const int64_t ks = a.nk_stride;
    const int lane = __lane_id();
    const int wave = static_cast<int>(blk * (kAdaptiveBlock / 64) + (threadIdx.x >> 6));
    const int64_t kbase = static_cast<int64_t>(wave) * 64 * KPT + lane;
    R *st = static_cast<R *>(a.state);

    // FP32 state: the propagated variable is delta = s + dh(k + A(t)), the
    // deviation from the instantaneous ground state -dh (see the FP32 notes
    // above); FP64 copies of the k phases give dh and the observables in FP64.
    constexpr bool kDelta = kComp;
    KPoint<R> p[KPT];
    // delta form: d(k + A) in FP64 from FP64 phases and laser rows, rounded to
    // FP32 per stage. FP32 phases or rows would carry a rounding error that is
    // fixed per k, or shared by all k at one time, and so shift the dynamical
    // phases coherently; rounding d itself leaves errors that average out.
    // ph: cos a_0..a_2, cos b_0..b_2, then the sines
    double ph[kDelta ? KPT : 1][12];
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = kbase + 64 * q;
        const double kx = a.kpts[k], ky = a.kpts[ks + k];
        p[q].w = a.weight[k];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double sa, ca, sb, cb;
            sincos(kx * a.va[j][0] + ky * a.va[j][1], &sa, &ca);
            sincos(kx * a.vb[j][0] + ky * a.vb[j][1], &sb, &cb);
            p[q].er[j] = static_cast<R>(ca);
            p[q].ei[j] = static_cast<R>(sa);
            p[q].er[j + 3] = static_cast<R>(cb);
            p[q].ei[j + 3] = static_cast<R>(sb);
            if constexpr (kDelta) {
                ph[q][j] = ca;
                ph[q][6 + j] = sa;
                ph[q][j + 3] = cb;
                ph[q][6 + j + 3] = sb;
            }
        }
        p[q].n = st[k];
        p[q].sx = st[ks + k];
        p[q].sy = st[2 * ks + k];
        p[q].sz = st[3 * ks + k];
        p[q].cn = p[q].cx = p[q].cy = p[q].cz = R(0);
        if (kComp) {
            p[q].cn = st[4 * ks + k];
            p[q].cx = st[5 * ks + k];
            p[q].cy = st[6 * ks + k];
            p[q].cz = st[7 * ks + k];
        }
    }
    const R c3 = static_cast<R>(a.c3), M0 = static_cast<R>(a.M0);
    // T1 is off on this path (the host rejects it): g1 = 0
    const R gh2 = static_cast<R>(0.5 * a.gam2), gh12 = static_cast<R>(-0.5 * a.gam2);
    const R ii = a.inter_intra ? R(1) : R(0);
    // this lane writes value (lane >> 3) of the wave's partial row of each sample
    double *part = a.partials + static_cast<int64_t>(wave) * kHaldaneNV + (lane >> 3);
    const int64_t row_stride = static_cast<int64_t>(a.nwaves) * kHaldaneNV;

    __shared__ R lds_rows[kAdaptiveBlock / 64][6 * kRow];
    R *rows = lds_rows[threadIdx.x >> 6];
    const LaserLane L = laser_lane(A);
    // eps of the reference vector: 2 % of a bound on |d| over the zone, so
    // that its rate stays below |dd/dt| / eps where the gap closes
    const double dmax = 3.0 * fabs(a.t1) + fabs(a.M0) + 3.0 * fabs(a.c3);
    const double eps2 = (0.02 * dmax) * (0.02 * dmax);
    __shared__ double lds_row64[kAdaptiveBlock / 64][kDelta ? 6 * kRow : 1];
    double *row64 = kDelta ? lds_row64[threadIdx.x >> 6] : nullptr;
    const double *end64 = row64 + 5 * kRow;  // the step end (window start: its time)
    // FP64 dh and its rate at the step end
    auto dh64 = [&](int q, Vec3<double> &dh, Vec3<double> &dht) {
        Vec3<double> d, td;
        d_and_rate<double, double>(ph[q], ph[q] + 6, end64, a, d, td);
        unit_and_rate(d, td, eps2, dh, dht);
    };
    // FP64 observables of the delta state at the step end: s = delta - dh,
    // ds/dt = 2 hf (the stage without the dh rate)
    auto observe64 = [&](int q, const Vec3<R> &hf, double o[8], double od[8]) {
        Vec3<double> dh, dht;
        dh64(q, dh, dht);
        KPoint<double> k64;
        for (int j = 0; j < 6; ++j) {
            k64.er[j] = ph[q][j];
            k64.ei[j] = ph[q][6 + j];
        }
        k64.w = p[q].w;
        const Vec3<double> s64{static_cast<double>(p[q].sx) - static_cast<double>(p[q].cx) - dh.x,
                               static_cast<double>(p[q].sy) - static_cast<double>(p[q].cy) - dh.y,
                               static_cast<double>(p[q].sz) - static_cast<double>(p[q].cz) - dh.z};
        const Vec3<double> f64{2.0 * hf.x, 2.0 * hf.y, 2.0 * hf.z};
        observe<double>(k64, static_cast<double>(p[q].n) - static_cast<double>(p[q].cn), s64, f64, end64, a,
                        a.inter_intra ? 1.0 : 0.0, o, od);
    };
    // stage of the delta form at the FP64 row r64: hf = d x y - dephasing
    // (f/2 of s), k = hf + dh_t / 2
    auto stage_delta = [&](int q, const double *r64, const Vec3<R> &y, Vec3<R> &hf, Vec3<R> &dh) {
        Vec3<double> d64, td64;
        d_and_rate<double, double>(ph[q], ph[q] + 6, r64, a, d64, td64);
        const Vec3<R> d{static_cast<R>(d64.x), static_cast<R>(d64.y), static_cast<R>(d64.z)};
        const Vec3<R> td{static_cast<R>(td64.x), static_cast<R>(td64.y), static_cast<R>(td64.z)};
        Vec3<R> dht;
        unit_and_rate(d, td, static_cast<R>(eps2), dh, dht);
        hf = half_f<R, DEPH>(d, y, gh2, gh12);
        return Vec3<R>{hf.x + R(0.5) * dht.x, hf.y + R(0.5) * dht.y, hf.z + R(0.5) * dht.z};
    };
    // Kahan pair (hi, c) holding v = hi - c from an FP64 value
    auto split = [](double v, R &hi, R &c) {
        hi = static_cast<R>(v);
        c = static_cast<R>(static_cast<double>(hi) - v);
    };

    int64_t m = out0;
    double t = A.atmin + A.dt_out * static_cast<double>(m);
    const double t_end = A.atmin + A.dt_out * static_cast<double>(out1);
    const int64_t max_iter = 64 * (out1 - out0) + 1024;  // step attempts allowed in this launch
    double h = uniform(A.hstep[wave]);
    int64_t nacc = 0, nrej = 0;
    bool failed = false;

    // Step ends carry the wave sums of the observables (this lane: value
    // lane >> 3) into the six-node ring; samples are its Lagrange
    // interpolants. At the window start: the sample itself, and the first
    // stage.
    Vec3<R> k1[KPT];
    Vec3<R> dh0[KPT];  // delta form: dh at the current step start (error scale)
    __shared__ double lds_ring[kAdaptiveBlock / 64][6];
    SampleRing ring;
    ring.cnt = 0;
    ring.t = lds_ring[threadIdx.x >> 6];
#pragma unroll
    for (int i = 0; i < 6; ++i) ring.D[i] = 0.0;
    for (int i = 0; i < 6; ++i) ring.t[i] = 0.0;
    auto emit = [&](double t_hi) {  // nt < 2^31 (host check): exact conversions
        m = ring.emit(static_cast<int>(m), static_cast<int>(out1), A.atmin, A.dt_out, t_hi, part, out0, row_stride);
    };
    {
        laser_rows<R>(L, t, 0.0, t, rows, row64);
        const R *row = rows + 5 * kRow;
        double o[8] = {0, 0, 0, 0, 0, 0, 0, 0}, od[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            if constexpr (kDelta) {
                Vec3<double> dh, dht;  // s -> delta = s + dh at the window start, in FP64
                dh64(q, dh, dht);
                split(static_cast<double>(p[q].sx) - static_cast<double>(p[q].cx) + dh.x, p[q].sx, p[q].cx);
                split(static_cast<double>(p[q].sy) - static_cast<double>(p[q].cy) + dh.y, p[q].sy, p[q].cy);
                split(static_cast<double>(p[q].sz) - static_cast<double>(p[q].cz) + dh.z, p[q].sz, p[q].cz);
                Vec3<R> hf;
                k1[q] = stage_delta(q, end64, Vec3<R>{p[q].sx, p[q].sy, p[q].sz}, hf, dh0[q]);
                observe64(q, hf, o, od);
            } else {
                k1[q] = half_f<R, DEPH>(stage_d(p[q], row, c3, M0), Vec3<R>{p[q].sx, p[q].sy, p[q].sz}, gh2, gh12);
                observe<R>(p[q], p[q].n - p[q].cn, Vec3<R>{p[q].sx - p[q].cx, p[q].sy - p[q].cy, p[q].sz - p[q].cz},
                           Vec3<R>{2 * k1[q].x, 2 * k1[q].y, 2 * k1[q].z}, row, a, ii, o, od);
            }
        }
        const double O0 = wave_reduce8(o);
        part[(m - out0) * row_stride] = O0;
        ring.push(t, O0);
        ++m;
    }

    int64_t iter = 0;
    while (t < t_end) {
        if (++iter > max_iter || !(h >= A.hmin)) {
            failed = true;
            break;
        }
        double hh = fmin(h, A.hmax);
        const bool last = t + 1.0001 * hh >= t_end;
        if (last) hh = t_end - t;
        const double tn = last ? t_end : t + hh;
        laser_rows<R>(L, t, hh, tn, rows, row64);

        Vec3<R> inc[KPT], y5[KPT], kv[KPT][6], dh1[KPT];
        double err_k = 0.0;
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            // stages 2..6 (stage 1 is k1)
            kv[q][0] = k1[q];
            const Vec3<R> y0{p[q].sx, p[q].sy, p[q].sz};
#pragma unroll
            for (int i = 1; i < 6; ++i) {
                Vec3<R> y = y0;
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    if (CK::a[i][j] != 0.0) {
                        const R aij = static_cast<R>(2.0 * CK::a[i][j] * hh);
                        y.x += aij * kv[q][j].x;
                        y.y += aij * kv[q][j].y;
                        y.z += aij * kv[q][j].z;
                    }
                }
                if constexpr (kDelta) {
                    Vec3<R> hf, dh;
                    kv[q][i] = stage_delta(q, row64 + (i - 1) * kRow, y, hf, dh);
                    if (i == 4) dh1[q] = dh;  // c = 1: the step end
                } else {
                    kv[q][i] = half_f<R, DEPH>(stage_d(p[q], rows + (i - 1) * kRow, c3, M0), y, gh2, gh12);
                }
            }
            // 5th-order increment (state precision) and the error estimate (FP64)
            inc[q] = Vec3<R>{0, 0, 0};
            double ex = 0, ey = 0, ez = 0;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                if (CK::b[i] != 0.0) {
                    const R bi = static_cast<R>(2.0 * CK::b[i] * hh);
                    inc[q].x += bi * kv[q][i].x;
                    inc[q].y += bi * kv[q][i].y;
                    inc[q].z += bi * kv[q][i].z;
                }
                if (CK::e[i] != 0.0) {
                    const double ei = 2.0 * CK::e[i] * hh;
                    ex += ei * static_cast<double>(kv[q][i].x);
                    ey += ei * static_cast<double>(kv[q][i].y);
                    ez += ei * static_cast<double>(kv[q][i].z);
                }
            }
            y5[q] = Vec3<R>{y0.x + inc[q].x, y0.y + inc[q].y, y0.z + inc[q].z};
            auto ratio = [&](double e, R y, R z) {  // 0 when the component and its error vanish (atol = 0)
                const double sc = A.atol + A.rtol * fmax(fabs(static_cast<double>(y)), fabs(static_cast<double>(z)));
                return e == 0.0 ? 0.0 : fabs(e) * recip(sc);
            };
            if constexpr (kDelta) {  // scales on s = delta - dh at both ends
                const Vec3<R> s0{y0.x - dh0[q].x, y0.y - dh0[q].y, y0.z - dh0[q].z};
                const Vec3<R> s1{y5[q].x - dh1[q].x, y5[q].y - dh1[q].y, y5[q].z - dh1[q].z};
                err_k = nanmax(err_k, nanmax(ratio(ex, s0.x, s1.x), nanmax(ratio(ey, s0.y, s1.y), ratio(ez, s0.z, s1.z))));
            } else {
                err_k = nanmax(err_k, nanmax(ratio(ex, y0.x, y5[q].x),
                                             nanmax(ratio(ey, y0.y, y5[q].y), ratio(ez, y0.z, y5[q].z))));
            }
        }
        const double err = wave_max(err_k);

        if (err <= 1.0) {
            const R *row7 = rows + 5 * kRow;
            double o[8] = {0, 0, 0, 0, 0, 0, 0, 0}, od[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
                Vec3<R> hf;
                if constexpr (kDelta) k1[q] = stage_delta(q, end64, y5[q], hf, dh0[q]);  // f(t+h, y5)
                else k1[q] = half_f<R, DEPH>(stage_d(p[q], row7, c3, M0), y5[q], gh2, gh12);
                if (kComp) {
                    kahan(p[q].sx, p[q].cx, inc[q].x);
                    kahan(p[q].sy, p[q].cy, inc[q].y);
                    kahan(p[q].sz, p[q].cz, inc[q].z);
                } else {
                    p[q].sx += inc[q].x;
                    p[q].sy += inc[q].y;
                    p[q].sz += inc[q].z;
                }
                if constexpr (kDelta) observe64(q, hf, o, od);
                else observe<R>(p[q], p[q].n - p[q].cn, Vec3<R>{p[q].sx - p[q].cx, p[q].sy - p[q].cy, p[q].sz - p[q].cz},
                                Vec3<R>{2 * k1[q].x, 2 * k1[q].y, 2 * k1[q].z}, row7, a, ii, o, od);
            }
            ring.push(tn, wave_reduce8(o));
            if (ring.cnt >= 6) emit(ring.t[2]);  // samples up to node c - 2
            t = tn;
            ++nacc;
            const double fac = err > 0.0 ? fmin(5.0, fmax(0.2, 0.9 * pow(err, -0.2))) : 5.0;
            if (!last) h = hh * fac;  // a shortened last step keeps the proposal it was cut from
        } else {
            h = hh * fmax(0.2, 0.9 * pow(err, -0.2));  // NaN err -> 0.2: shrinks to hmin and fails
            ++nrej;
        }
    }

    if (!failed) emit(t_end);  // the window end: every pending sample
    if constexpr (kDelta) {  // delta -> s at the window end (row64 holds the last step end, t_end)
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            Vec3<double> dh, dht;
            dh64(q, dh, dht);
            split(static_cast<double>(p[q].sx) - static_cast<double>(p[q].cx) - dh.x, p[q].sx, p[q].cx);
            split(static_cast<double>(p[q].sy) - static_cast<double>(p[q].cy) - dh.y, p[q].sy, p[q].cy);
            split(static_cast<double>(p[q].sz) - static_cast<double>(p[q].cz) - dh.z, p[q].sz, p[q].cz);
        }
    }
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = kbase + 64 * q;
        st[k] = p[q].n;
        st[ks + k] = p[q].sx;
        st[2 * ks + k] = p[q].sy;
        st[3 * ks + k] = p[q].sz;
        if (kComp) {
            st[4 * ks + k] = p[q].cn;
            st[5 * ks + k] = p[q].cx;
            st[6 * ks + k] = p[q].cy;
            st[7 * ks + k] = p[q].cz;
        }
    }
    if (lane == 0) {
        A.hstep[wave] = h;
        A.counts[2 * wave] += nacc;
        A.counts[2 * wave + 1] += nrej;
        if (failed) *A.fail = 1;
    }
}

// f/2 of the Bloch equation at d (1/|d|^2 = inv): d x y - (g2/2) y + ((g2 - g1)/2) (y.d/|d|^2) d
// d(k + A) and 1/|d|^2 at a stage time: the sums of
// phases_at, each component formed as one FMA chain (a stage needs no single
// phase, only d)
__device__ __forceinline__ void stage_d_inv(const KPoint<double> &p, const double *row, double c3, double M0,
                                            Vec3<double> &d, double &inv, double &dd)
{
    double dx = p.er[0] * row[0], dy = p.er[0] * row[1], sb = p.er[3] * row[7];
    dx = fma(-p.ei[0], row[1], dx);
    dy = fma(p.ei[0], row[0], dy);
    sb = fma(p.ei[3], row[6], sb);
#pragma unroll
    for (int j = 1; j < 3; ++j) {
        dx = fma(p.er[j], row[2 * j], dx);
        dx = fma(-p.ei[j], row[2 * j + 1], dx);
        dy = fma(p.er[j], row[2 * j + 1], dy);
        dy = fma(p.ei[j], row[2 * j], dy);
        sb = fma(p.er[j + 3], row[7 + 2 * j], sb);
        sb = fma(p.ei[j + 3], row[6 + 2 * j], sb);
    }
    d = Vec3<double>{dx, dy, fma(-c3, sb, M0)};
    dd = fma(d.x, d.x, fma(d.y, d.y, d.z * d.z));
    // d = 0 (a gap closing) gives a finite inv and a zero dephasing projection
    inv = rcp_nr(fmax(dd, 1e-300));
}

// The single phases t1 e^{i(k+A).a_j} and e^{i(k+A).b_j} of a row (the
// products of phases_at, without the sums)
__device__ __forceinline__ void step_end_phases(const KPoint<double> &p, const double *row, Phases<double> &e)
{
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        e.ar[j] = fma(p.er[j], row[2 * j], -(p.ei[j] * row[2 * j + 1]));
        e.ai[j] = fma(p.er[j], row[2 * j + 1], p.ei[j] * row[2 * j]);
        e.br[j] = fma(p.er[j + 3], row[6 + 2 * j], -(p.ei[j + 3] * row[7 + 2 * j]));
        e.bi[j] = fma(p.er[j + 3], row[7 + 2 * j], p.ei[j + 3] * row[6 + 2 * j]);
    }
}

template <bool DEPH>
__device__ __forceinline__ Vec3<double> half_f_at(const Vec3<double> &d, double inv, const Vec3<double> &y, double gh2,
                                                  double gh12)
{
    Vec3<double> f;
    f.x = d.y * y.z - d.z * y.y;
    f.y = d.z * y.x - d.x * y.z;
    f.z = d.x * y.y - d.y * y.x;
    if (DEPH) {
        const double c = -gh12 * (y.x * d.x + y.y * d.y + y.z * d.z) * inv;
        f.x += c * d.x - gh2 * y.x;
        f.y += c * d.y - gh2 * y.y;
        f.z += c * d.z - gh2 * y.z;
    }
    return f;
}

// The FP64 stepper. Per accepted step of a wave: five new stage times (the
// stage at c = 1 is the step end, whose phases and d also give the next
// step's first stage and the observables), the FP64 error estimate and one
// 8-value wave reduction of the observables at the step end. Samples are
// written two step ends later, from the six-node ring (centred stencil;
// one-sided at the window ends, where all pending samples are written).
template <bool DEPH, int KPT>
__device__ void adaptive_body64(const AdaptiveArgs &A, int64_t blk, const int64_t out0, const int64_t out1)
{
    using CK = CashKarp;
    const HaldaneArgs &a = A.h;
    const int64_t ks = a.nk_stride;
    const int lane = __lane_id();
    const int wave = static_cast<int>(blk * (kAdaptiveBlock / 64) + (threadIdx.x >> 6));
    const int64_t kbase = static_cast<int64_t>(wave) * 64 * KPT + lane;
    double *st = static_cast<double *>(a.state);

    KPoint<double> p[KPT];
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = kbase + 64 * q;
        const double kx = a.kpts[k], ky = a.kpts[ks + k];
        p[q].w = a.weight[k];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            sincos(kx * a.va[j][0] + ky * a.va[j][1], &p[q].ei[j], &p[q].er[j]);
            sincos(kx * a.vb[j][0] + ky * a.vb[j][1], &p[q].ei[j + 3], &p[q].er[j + 3]);
        }
        p[q].n = st[k];
        p[q].sx = st[ks + k];
        p[q].sy = st[2 * ks + k];
        p[q].sz = st[3 * ks + k];
    }
    const double c3 = a.c3, M0 = a.M0;
    const double gh2 = 0.5 * a.gam2, gh12 = -0.5 * a.gam2;  // T1 off on this path
    const double ii = a.inter_intra ? 1.0 : 0.0;
    double *part = a.partials + static_cast<int64_t>(wave) * kHaldaneNV + (lane >> 3);
    const int64_t row_stride = static_cast<int64_t>(a.nwaves) * kHaldaneNV;
    __shared__ double lds_rows[kAdaptiveBlock / 64][6 * kRow];
    double *rows = lds_rows[threadIdx.x >> 6];
    const LaserLane L = laser_lane(A);

    // the arguments used in the loops, in registers (stores through part could alias A)
    const double atmin = A.atmin, dt_out = A.dt_out, atol = A.atol, rtol = A.rtol, hmax = A.hmax, hmin = A.hmin;
    int64_t m = out0;
    double t = atmin + dt_out * static_cast<double>(m);
    const double t_end = atmin + dt_out * static_cast<double>(out1);
    const int64_t max_iter = 64 * (out1 - out0) + 1024;
    double h = uniform(A.hstep[wave]);
    int64_t nacc = 0, nrej = 0;
    bool failed = false;
    __shared__ double lds_ring[kAdaptiveBlock / 64][6];
    SampleRing ring;
    ring.cnt = 0;
    ring.t = lds_ring[threadIdx.x >> 6];
#pragma unroll
    for (int i = 0; i < 6; ++i) ring.D[i] = 0.0;
    for (int i = 0; i < 6; ++i) ring.t[i] = 0.0;
    // samples up to t_hi from the ring
    auto emit = [&](double t_hi) {  // nt < 2^31 (host check): exact conversions
        m = ring.emit(static_cast<int>(m), static_cast<int>(out1), atmin, dt_out, t_hi, part, out0, row_stride);
    };

    Vec3<double> k1[KPT];
    {  // the window start: its sample, the first node and the first stage
        laser_rows<double>(L, t, 0.0, t, rows);
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            const Phases<double> e = phases_at(p[q], rows + 5 * kRow, c3, M0);
            const Vec3<double> y{p[q].sx, p[q].sy, p[q].sz};
            k1[q] = half_f_at<DEPH>(e.d, e.inv, y, gh2, gh12);
            sample_state(e, p[q].w, p[q].n, y.x, y.y, y.z, a, ii, v);
        }
        const double val = wave_reduce8(v);
        part[0] = val;
        ring.push(t, val);
        ++m;
    }

    int64_t iter = 0;
    while (t < t_end) {
        if (++iter > max_iter || !(h >= hmin)) {
            failed = true;
            break;
        }
        double hh = fmin(h, hmax);
        const bool last = t + 1.0001 * hh >= t_end;
        if (last) hh = t_end - t;
        const double tn = last ? t_end : t + hh;
        laser_rows<double>(L, t, hh, tn, rows);

        Vec3<double> inc[KPT];
        Vec3<double> d_end[KPT];  // d, 1/|d|^2 and |d|^2 at the step end (the stage at c = 1)
        double inv_end[KPT], dd_end[KPT];
        double err_k = 0.0;
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            Vec3<double> kv[6];
            kv[0] = k1[q];
            const Vec3<double> y0{p[q].sx, p[q].sy, p[q].sz};
#pragma unroll
            for (int i = 1; i < 6; ++i) {
                Vec3<double> y = y0;
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    if (CK::a[i][j] != 0.0) {
                        const double aij = 2.0 * CK::a[i][j] * hh;
                        y.x = fma(aij, kv[j].x, y.x);
                        y.y = fma(aij, kv[j].y, y.y);
                        y.z = fma(aij, kv[j].z, y.z);
                    }
                }
                Vec3<double> d;
                double inv, dd;
                stage_d_inv(p[q], rows + (i - 1) * kRow, c3, M0, d, inv, dd);
                kv[i] = half_f_at<DEPH>(d, inv, y, gh2, gh12);
                if (CK::c[i] == 1.0) {
                    d_end[q] = d;
                    inv_end[q] = inv;
                    dd_end[q] = dd;
                }
            }
            inc[q] = Vec3<double>{0, 0, 0};
            double ex = 0, ey = 0, ez = 0;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                if (CK::b[i] != 0.0) {
                    const double bi = 2.0 * CK::b[i] * hh;
                    inc[q].x = fma(bi, kv[i].x, inc[q].x);
                    inc[q].y = fma(bi, kv[i].y, inc[q].y);
                    inc[q].z = fma(bi, kv[i].z, inc[q].z);
                }
                if (CK::e[i] != 0.0) {
                    const double ei = 2.0 * CK::e[i] * hh;
                    ex = fma(ei, kv[i].x, ex);
                    ey = fma(ei, kv[i].y, ey);
                    ez = fma(ei, kv[i].z, ez);
                }
            }
            // |e| / scale; a zero scale (atol = 0 with a vanishing component and increment)
            // counts as 1e-300, so a vanishing error still gives 0
            auto ratio = [&](double e, double y, double z) {
                const double sc = atol + rtol * fmax(fabs(y), fabs(y + z));
                return fabs(e) * recip(fmax(sc, 1e-300));
            };
            const double rx = ratio(ex, y0.x, inc[q].x), ry = ratio(ey, y0.y, inc[q].y),
                         rz = ratio(ez, y0.z, inc[q].z);
            // a NaN ratio (a non-finite state) counts as +inf: rejected, the step shrinks to
            // hmin and the solve fails, as for NaN; the maxima then need no NaN test
            const double sum = rx + ry + rz;
            err_k = fmax(err_k, sum == sum ? fmax(fmax(rx, ry), rz) : __builtin_inf());
        }
        const double err = wave_max_nonneg(err_k);

        if (err <= 1.0) {
            double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < KPT; ++q) {
                p[q].sx += inc[q].x;
                p[q].sy += inc[q].y;
                p[q].sz += inc[q].z;
                const Vec3<double> y{p[q].sx, p[q].sy, p[q].sz};
                // the single phases at the step end from its row (the stage at c = 1 formed
                // only their sums: d, |d|^2 and 1/|d|^2, kept from there)
                Phases<double> e1;
                step_end_phases(p[q], rows + 5 * kRow, e1);
                e1.d = d_end[q];
                e1.dd = dd_end[q];
                e1.inv = inv_end[q];
                k1[q] = half_f_at<DEPH>(e1.d, e1.inv, y, gh2, gh12);  // f(t + h, y5)
                sample_state(e1, p[q].w, p[q].n, y.x, y.y, y.z, a, ii, v);
            }
            ring.push(tn, wave_reduce8(v));
            if (ring.cnt >= 6) emit(ring.t[2]);  // samples up to node c - 2
            t = tn;
            ++nacc;
            const double fac = err > 0.0 ? fmin(5.0, fmax(0.2, 0.9 * inv_fifth_root(err))) : 5.0;
            if (!last) h = hh * fac;  // a shortened last step keeps the proposal it was cut from
        } else {
            h = hh * fmax(0.2, 0.9 * inv_fifth_root(err));  // NaN err -> 0.2: shrinks to hmin and fails
            ++nrej;
        }
    }
    if (!failed) emit(t_end);  // the window end: every pending sample

#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int64_t k = kbase + 64 * q;
        st[k] = p[q].n;
        st[ks + k] = p[q].sx;
        st[2 * ks + k] = p[q].sy;
        st[3 * ks + k] = p[q].sz;
    }
    if (lane == 0) {
        A.hstep[wave] = h;
        A.counts[2 * wave] += nacc;
        A.counts[2 * wave + 1] += nrej;
        if (failed) *A.fail = 1;
    }
}

// Block b works for solve block_inst[b] on that solve's block
// b - block_first[solve] (a single solve: block_inst null, solve 0). The
// window is the next nwin samples after the solve's cursor; solves with none
// left leave at once. One kernel serves both cases, so a solve gives the
// same bits alone and in a batch.
template <typename R, bool DEPH, int KPT>
__global__ __launch_bounds__(kAdaptiveBlock)
__attribute__((amdgpu_waves_per_eu(KPT == 1 && sizeof(R) == 8 ? 3 : 1))) void haldane_adaptive_kernel(const AdaptiveArgs *__restrict__ insts,
                                                                  const int32_t *__restrict__ block_inst,
                                                                  const int32_t *__restrict__ block_first,
                                                                  int64_t nwin)
{
    const int inst = block_inst ? block_inst[blockIdx.x] : 0;
    const AdaptiveArgs &A = insts[inst];
    const int64_t out0 = *A.cursor;
    const int64_t out1 = min(out0 + nwin, A.nt);
    if (out1 <= out0) return;  // whole block
    const int64_t blk = static_cast<int64_t>(blockIdx.x) - (block_inst ? block_first[inst] : 0);
    if constexpr (sizeof(R) == 8) adaptive_body64<DEPH, KPT>(A, blk, out0, out1);
    else adaptive_body<R, DEPH, KPT>(A, blk, out0, out1);
}

__global__ void adaptive_advance_kernel(const AdaptiveArgs *__restrict__ insts, int n, int64_t nwin)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = *insts[i].cursor;
    *insts[i].cursor = min(c + nwin, insts[i].nt);
}

template <typename R, int KPT>
void launch_adaptive(const AdaptiveArgs *d, const int32_t *bi, const int32_t *bf, int nblocks, int64_t nwin, bool deph,
                     hipStream_t s)
{
    if (deph) haldane_adaptive_kernel<R, true, KPT><<<nblocks, kAdaptiveBlock, 0, s>>>(d, bi, bf, nwin);
    else haldane_adaptive_kernel<R, false, KPT><<<nblocks, kAdaptiveBlock, 0, s>>>(d, bi, bf, nwin);
}

}  // namespace

hipError_t launch_haldane_adaptive(const AdaptiveArgs *d_insts, const int32_t *d_block_inst,
                                   const int32_t *d_block_first, int nblocks, int64_t nwin, int precision,
                                   bool deph, int kpt, hipStream_t s)
{
    if (nblocks <= 0 || nwin <= 0) return hipSuccess;
    const bool f32 = precision == SBE_PREC_FP32;
    if (kpt == 2) {
        if (f32) launch_adaptive<float, 2>(d_insts, d_block_inst, d_block_first, nblocks, nwin, deph, s);
        else launch_adaptive<double, 2>(d_insts, d_block_inst, d_block_first, nblocks, nwin, deph, s);
    } else {
        if (f32) launch_adaptive<float, 1>(d_insts, d_block_inst, d_block_first, nblocks, nwin, deph, s);
        else launch_adaptive<double, 1>(d_insts, d_block_inst, d_block_first, nblocks, nwin, deph, s);
    }
    return hipGetLastError();
}

hipError_t launch_adaptive_advance(const AdaptiveArgs *d_insts, int n, int64_t nwin, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    adaptive_advance_kernel<<<(n + 63) / 64, 64, 0, s>>>(d_insts, n, nwin);
    return hipGetLastError();
}

}  // namespace sbe
