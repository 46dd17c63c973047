// hc_riemann.hpp -- normal (rpn2) and transverse (rpt2) Riemann solvers for
// one face, evaluated in registers by one lane.
//
// Each routine restates the reference Fortran solver for a single interface
// with the same operation order, so the FP64 instantiation (built without FMA
// contraction) reproduces the reference bit for bit:
//   Adv     applications/clawpack/advection/2d/all/clawpack46_rpn2adv.f:30-58,
//           clawpack46_rpt2adv.f:18-25
//   SWE     applications/clawpack/shallow/2d/rp/clawpack46_rpn2.f:83-245,
//           clawpack46_rpt2.f:34-72 (Roe averages shared with rpn2 as /comroe/ does)
//   FwCons  applications/clawpack/transport/2d/all/clawpack46_rpn2cons_fw_manifold.f:23-46,
//           clawpack46_rpt2cons_manifold.f:25-63 (with flux2fw's dsign(1, s) weight)
//   Euler5  applications/clawpack/euler/2d/rp/clawpack46_rpn2_euler5.f:62-278,
//           clawpack46_rpt2_euler5.f:48-118 (rpt2 recomputes the same Roe
//           averages from the same two cells; they are reused here)
//
// The Clawpack 5 solvers of the same applications (C5 = true) differ only in
// operation order, which the C5 instantiations follow:
//   SWE     clawpack5_rpn2.f90:76,78 and clawpack5_rpt2.f90:59,61 scale by
//           (0.5/a) where the 4.6 solvers divide by 2a;
//   Euler5  clawpack5_rpt2_euler5.f90:67 takes gamma1/c**2 (c = sqrt of the
//           4.6 a2) for the transverse split; its rpn2 and the Adv pair
//           compute exactly what the 4.6 ones do.
//
// Naming: qa is the cell on the low side of the face ("qr(i-1)" in the
// reference), qb the cell on the high side ("ql(i)").  DIR = 0 for an x-face,
// 1 for a y-face (reference ixy = DIR+1).
#pragma once
#include "hc_common.hpp"

namespace hc {

enum { SOLVER_ADV = 0, SOLVER_SWE = 1, SOLVER_EULER5 = 2, SOLVER_EULER_HLLE = 3, SOLVER_TRANSPORT_FW = 4 };

template <class R, int ME_, int MW_> struct FaceT {
    static constexpr int ME = ME_, MW = MW_;
    R w[MW][ME];
    R s[MW];
    R am[ME];
    R ap[ME];
};

struct PhysParams {
    double grav;
    double gamma;
    int efix;
    double sgrav;  // sqrt(grav), for the FP32 shallow-water solver
};

// ---------------------------------------------------------------------------
// Scalar advection with edge velocities (aux field 0 = u at the left edge,
// field 1 = v at the bottom edge).
// ---------------------------------------------------------------------------
template <class R> struct Adv {
    static constexpr int ME = 1, MW = 1, MAUX = 2;
    static constexpr bool HAS_CONS = false;
    static constexpr bool FW = false;  // f-wave edge-velocity form (FwCons)
    using Face = FaceT<R, 1, 1>;
    // transverse speeds for the amdq split (cell on the low side) and the
    // apdq split (cell on the high side): {down, up}
    struct Trans { R m_dn, m_up, p_dn, p_up; };

    static constexpr int NC = 0;  // no per-cell data
    __device__ static inline void cell(const R*, const PhysParams&, R*) {}
    // component m of wave w is zero by construction (none here)
    template <int DIR> __device__ static constexpr bool wz(int, int) { return false; }
    static constexpr bool RPT2_SCALED = false;
    static constexpr bool SMAX = false;
    __device__ static inline R smax(const Face&, const Trans&) { return R(0); }
    template <int DIR> __device__ static void rpt2_scaled(const R*, const Trans&, R, R*, R*) {}

    template <int DIR>
    __device__ static inline void rpn2(const R* qa, const R* qb, const R*, const R*, R sface,
                                       const PhysParams&, Face& f, Trans&)
    {
        R delta = qb[0] - qa[0];
        f.w[0][0] = delta;
        f.s[0] = sface;
        f.am[0] = R(0) + zmin(f.s[0]) * f.w[0][0];
        f.ap[0] = R(0) + zmax(f.s[0]) * f.w[0][0];
    }

    template <int DIR>
    __device__ static inline void rpt2(int imp, const R* asdq, const Trans& t,
                                       const PhysParams&, R* bm, R* bp)
    {
        R dn = imp == 1 ? t.m_dn : t.p_dn;
        R up = imp == 1 ? t.m_up : t.p_up;
        bm[0] = zmin(dn) * asdq[0];
        bp[0] = zmax(up) * asdq[0];
    }
};

// ---------------------------------------------------------------------------
// Conservative transport q_t + (u q)_x + (v q)_y = 0 in f-wave form, with the
// edge data of the transport applications (aux components, 0-based: 1 / 2
// normal velocity on a cell's left / right edge, 3 / 4 on its bottom / top
// edge, 5 / 6 the x- / y-edge length ratio; component 0, the capacity, is not
// read).  The single wave is the flux jump (the f-wave) and its speed the mean
// of the two edge velocities; the kernel forms the velocities of a face from
// the two cells' edge data (rpn2f / mean) and takes the second-order weight
// sign(s) instead of |s| (fw_weight).
// ---------------------------------------------------------------------------
template <class R> struct FwCons {
    static constexpr int ME = 1, MW = 1, MAUX = 7;
    static constexpr bool HAS_CONS = false;
    static constexpr bool FW = true;
    using Face = FaceT<R, 1, 1>;
    // mean transverse velocities for the amdq split (edge length of the
    // low-side cell) and the apdq split (of the high-side cell): {down, up}
    struct Trans { R m_dn, m_up, p_dn, p_up; };

    static constexpr int NC = 0;
    __device__ static inline void cell(const R*, const PhysParams&, R*) {}
    template <int DIR> __device__ static constexpr bool wz(int, int) { return false; }
    static constexpr bool RPT2_SCALED = false;
    static constexpr bool SMAX = false;
    __device__ static inline R smax(const Face&, const Trans&) { return R(0); }
    template <int DIR> __device__ static void rpt2_scaled(const R*, const Trans&, R, R*, R*) {}

    // face between qa (low side) and qb: g the edge length of the high-side
    // cell, uhi its low-edge velocity, ulo the low-side cell's high-edge
    // velocity (rpn2cons_fw_manifold.f:29-44)
    __device__ static inline void rpn2f(R qa, R qb, R g, R uhi, R ulo, Face& f)
    {
        const R ur = g * uhi;
        const R ul = g * ulo;
        const R fw = ur * qb - ul * qa;
        const R uhat = (ul + ur) * R(0.5);
        f.w[0][0] = fw;
        f.s[0] = uhat;
        f.am[0] = uhat >= R(0) ? R(0) : fw;
        f.ap[0] = uhat >= R(0) ? fw : R(0);
    }

    // transverse mean velocity on one side (rpt2cons_manifold.f:33-36,45-48)
    __device__ static inline R mean(R g, R vhi, R vlo) { return (g * vhi + g * vlo) * R(0.5); }

    // second-order weight of a wave of speed s: dsign(1, s) (flux2fw.f:176)
    __device__ static inline R fw_weight(R s)
    {
        if constexpr (sizeof(R) == 8) return __builtin_copysign(1.0, s);
        else return __builtin_copysignf(1.0f, s);
    }

    template <int DIR>
    __device__ static inline void rpt2(int imp, const R* asdq, const Trans& t,
                                       const PhysParams&, R* bm, R* bp)
    {
        R dn = imp == 1 ? t.m_dn : t.p_dn;
        R up = imp == 1 ? t.m_up : t.p_up;
        bm[0] = zmin(dn) * asdq[0];
        bp[0] = zmax(up) * asdq[0];
    }
};

// ---------------------------------------------------------------------------
// Shallow water, Roe solver with Harten-Hyman entropy fix.
// ---------------------------------------------------------------------------
template <class R, bool C5> struct SweT {
    static constexpr int ME = 3, MW = 3, MAUX = 0;
    static constexpr bool HAS_CONS = true;
    static constexpr bool FW = false;  // f-wave edge-velocity form (FwCons)
    using Face = FaceT<R, 3, 3>;
    struct Trans { R u, v, a; };

    // Conservative flux f(q) in direction DIR, always in FP64, for the
    // time-sync edge registers (rpn2_cons_update of the radial-dam
    // application, applications/clawpack/shallow/2d/radialdam/rpn2cons_update.f:24-48).
    template <int DIR>
    __device__ static inline void cons_flux(const double* q, double grav, double* f)
    {
#pragma clang fp contract(off)
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        const double hun = q[mu], hut = q[mv], h = q[0];
        const double un = hun / h;
        f[0] = hun;
        f[mu] = hun * hun / h + 0.5 * grav * (h * h);
        f[mv] = un * hut;
    }

    // Per-cell data: sqrt(h), hu/sqrt(h), hv/sqrt(h), hu/h, hv/h, sqrt(g h),
    // the operations the reference evaluates per face side for the Roe
    // averages (clawpack46_rpn2.f:91-99) and the entropy-fix speeds
    // (:170-171, :207-208).  The FP64 kernels compute them once per cell and
    // share them between the four faces the cell touches (the same
    // operations, so no bit changes); the FP32 kernels (NC = 0) evaluate them
    // per face, which keeps their register footprint within two waves per SIMD.
    static constexpr int NC = sizeof(R) == 8 ? 6 : 0;
    // component m of wave w is zero by construction: the shear wave carries
    // only the transverse momentum
    template <int DIR> __device__ static constexpr bool wz(int w, int m)
    {
        return w == 1 && m != (DIR == 0 ? 2 : 1);
    }
    __device__ static inline void cell(const R* q, const PhysParams& pp, R* cd)
    {
        using MM = M<R>;
        const R g = R(pp.grav);
        const R sh = MM::sqrt_(q[0]);
        cd[0] = sh;
        cd[1] = MM::divp(q[1], sh);
        cd[2] = MM::divp(q[2], sh);
        cd[3] = MM::divp(q[1], q[0]);
        cd[4] = MM::divp(q[2], q[0]);
        cd[5] = MM::sqrt_(g * q[0]);
    }

    // FP32 form of rpn2: the same waves, speeds and entropy fix, evaluated
    // with fewer transcendental instructions, rounding aside.  One reciprocal
    // square root per state gives sqrt(h) = h r, hu/sqrt(h) = hu r,
    // u = hu r^2 and sqrt(g h) = sqrt(g) sqrt(h); one of g hbar gives both
    // a = g hbar t and 1/(2a) = t/2; the shear wave's two zero components are
    // left out of the sums; the transonic entropy-fix fractions divide by a
    // positive difference with the hardware reciprocal.
    template <int DIR>
    __device__ static inline void rpn2_f32(const R* qa, const R* qb, const PhysParams& pp, Face& f,
                                           Trans& roe)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        const R g = R(pp.grav), sg = R(pp.sgrav);
        const R ra = __builtin_amdgcn_rsqf(qa[0]), rb = __builtin_amdgcn_rsqf(qb[0]);
        const R hsl = qa[0] * ra, hsr = qb[0] * rb;
        const R iq = __builtin_amdgcn_rcpf(hsl + hsr);
        const R u = (qa[mu] * ra + qb[mu] * rb) * iq;
        const R v = (qa[mv] * ra + qb[mv] * rb) * iq;
        const R gh = g * ((qa[0] + qb[0]) * R(0.5));
        const R t = __builtin_amdgcn_rsqf(gh);
        const R a = gh * t, i2a = R(0.5) * t;
        roe.u = u; roe.v = v; roe.a = a;
        const R d0 = qb[0] - qa[0], d1 = qb[mu] - qa[mu], d2 = qb[mv] - qa[mv];
        const R a1 = ((u + a) * d0 - d1) * i2a;
        const R a3 = (d1 - (u - a) * d0) * i2a;
        const R a2 = d2 - v * d0;
        const R s0 = u - a, s2 = u + a;
        f.w[0][0] = a1; f.w[0][mu] = a1 * s0; f.w[0][mv] = a1 * v; f.s[0] = s0;
        f.w[1][0] = R(0); f.w[1][mu] = R(0); f.w[1][mv] = a2; f.s[1] = u;
        f.w[2][0] = a3; f.w[2][mu] = a3 * s2; f.w[2][mv] = a3 * v; f.s[2] = s2;
        if (!pp.efix) {
            const R m0 = zmin(s0), m2 = zmin(s2), m1 = zmin(u);
            const R p0 = zmax(s0), p2 = zmax(s2), p1 = zmax(u);
            for (int m = 0; m < 3; m++) {
                f.am[m] = m0 * f.w[0][m] + m2 * f.w[2][m];
                f.ap[m] = p0 * f.w[0][m] + p2 * f.w[2][m];
            }
            f.am[mv] = f.am[mv] + m1 * a2;
            f.ap[mv] = f.ap[mv] + p1 * a2;
            return;
        }
        // entropy fix (the branches of rpn2): amdq from the left-going parts
        R am0 = R(0), amu = R(0), amv = R(0);
        do {
            const R sa = qa[mu] * (ra * ra) - sg * hsl;  // u - c of the low state
            if (sa > R(0) && s0 > R(0)) break;
            const R h1 = qa[0] + a1, hu1 = qa[mu] + a1 * s0;
            const R r1 = __builtin_amdgcn_rsqf(h1);
            const R s1 = hu1 * (r1 * r1) - sg * (h1 * r1);
            R sfract;
            if (sa < R(0) && s1 > R(0))
                sfract = sa * ((s1 - s0) * __builtin_amdgcn_rcpf(s1 - sa));  // s1 - sa > 0
            else if (s0 < R(0))
                sfract = s0;
            else
                sfract = R(0);
            am0 = sfract * a1; amu = sfract * f.w[0][mu]; amv = sfract * f.w[0][mv];
            if (u > R(0)) break;
            amv = amv + u * a2;
            const R sb = qb[mu] * (rb * rb) + sg * hsr;  // u + c of the high state
            const R h3 = qb[0] - a3, hu3 = qb[mu] - a3 * s2;
            const R r3 = __builtin_amdgcn_rsqf(h3);
            const R s3 = hu3 * (r3 * r3) + sg * (h3 * r3);
            if (s3 < R(0) && sb > R(0))
                sfract = s3 * ((sb - s2) * __builtin_amdgcn_rcpf(sb - s3));  // sb - s3 > 0
            else if (s2 < R(0))
                sfract = s2;
            else
                break;
            am0 = am0 + sfract * a3; amu = amu + sfract * f.w[2][mu]; amv = amv + sfract * f.w[2][mv];
        } while (0);
        f.am[0] = am0; f.am[mu] = amu; f.am[mv] = amv;
        f.ap[0] = (s0 * a1 + s2 * a3) - am0;
        f.ap[mu] = (s0 * f.w[0][mu] + s2 * f.w[2][mu]) - amu;
        f.ap[mv] = (s0 * f.w[0][mv] + u * a2 + s2 * f.w[2][mv]) - amv;
    }

    // FP32 form of rpt2: bm = sum_w min(sb_w, 0) alpha_w r_w with the
    // eigenvectors r_1 = (1, u, v - a), r_2 = (0, 1, 0), r_3 = (1, u, v + a)
    // written out, so the zero components cost nothing.
    // FP32 split scaled by sc (the step's dt/(2 dx)): sc enters the three
    // wave coefficients instead of the six outputs
    static constexpr bool RPT2_SCALED = sizeof(R) == 4;
    // FP32: the largest |s| of a face, |u| + a (u the normal Roe velocity):
    // |fl(u - a)|, |u| and |fl(u + a)| are at most fl(|u| + a), and one of them
    // equals it, so dt/dx (|u| + a) is the maximum of dt/dx |s_w| bit for bit
    static constexpr bool SMAX = sizeof(R) == 4;
    __device__ static inline R smax(const Face&, const Trans& roe) { return M<R>::abs_(roe.u) + roe.a; }
    template <int DIR>
    __device__ static inline void rpt2_scaled(const R* asdq, const Trans& roe, R sc, R* bm, R* bp)
    {
        rpt2_f32<DIR>(asdq, roe, bm, bp, sc);
    }

    template <int DIR>
    __device__ static inline void rpt2_f32(const R* asdq, const Trans& roe, R* bm, R* bp, R sc = R(1))
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        const R u = roe.u, v = roe.v, a = roe.a;
        const R i2a = (R(0.5) * sc) * __builtin_amdgcn_rcpf(a);
        const R sm = v - a, sp = v + a;
        const R a1 = (sp * asdq[0] - asdq[mv]) * i2a;
        const R a3 = (asdq[mv] - sm * asdq[0]) * i2a;
        const R a2 = (asdq[mu] - u * asdq[0]) * sc;
        const R m1 = zmin(sm) * a1, m2 = zmin(v) * a2, m3 = zmin(sp) * a3;
        const R p1 = zmax(sm) * a1, p2 = zmax(v) * a2, p3 = zmax(sp) * a3;
        bm[0] = m1 + m3; bm[mu] = u * (m1 + m3) + m2; bm[mv] = sm * m1 + sp * m3;
        bp[0] = p1 + p3; bp[mu] = u * (p1 + p3) + p2; bp[mv] = sm * p1 + sp * p3;
    }

    template <int DIR>
    __device__ static inline void rpn2(const R* qa, const R* qb, const R* ca, const R* cb, R,
                                       const PhysParams& pp, Face& f, Trans& roe)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        if constexpr (sizeof(R) == 4) {
            rpn2_f32<DIR>(qa, qb, pp, f, roe);
            return;
        }
        using MM = M<R>;
        const R g = R(pp.grav);
        R h = (qa[0] + qb[0]) / R(2);
        R hsl, hsr, wua, wub, wva, wvb;
        if constexpr (NC > 0) {
            hsl = ca[0], hsr = cb[0];
            wua = ca[mu], wub = cb[mu], wva = ca[mv], wvb = cb[mv];
        } else {
            hsl = MM::sqrt_(qa[0]), hsr = MM::sqrt_(qb[0]);
            wua = MM::divp(qa[mu], hsl), wub = MM::divp(qb[mu], hsr);
            wva = MM::divp(qa[mv], hsl), wvb = MM::divp(qb[mv], hsr);
        }
        R hsq2 = hsl + hsr;
        R u = MM::divp(wua + wub, hsq2);
        R v = MM::divp(wva + wvb, hsq2);
        R a = MM::sqrt_(g * h);
        roe.u = u; roe.v = v; roe.a = a;

        R d0 = qb[0] - qa[0];
        R d1 = qb[mu] - qa[mu];
        R d2 = qb[mv] - qa[mv];
        R a1, a3;
        R a2 = -v * d0 + d2;
        if constexpr (C5) {
            const R ha = MM::divp(R(0.5), a);
            a1 = ((u + a) * d0 - d1) * ha;
            a3 = (-(u - a) * d0 + d1) * ha;
        } else {
            a1 = MM::divp((u + a) * d0 - d1, R(2) * a);
            a3 = MM::divp(-(u - a) * d0 + d1, R(2) * a);
        }

        f.w[0][0] = a1; f.w[0][mu] = a1 * (u - a); f.w[0][mv] = a1 * v;
        f.s[0] = u - a;
        f.w[1][0] = R(0); f.w[1][mu] = R(0); f.w[1][mv] = a2;
        f.s[1] = u;
        f.w[2][0] = a3; f.w[2][mu] = a3 * (u + a); f.w[2][mv] = a3 * v;
        f.s[2] = u + a;

        if (!pp.efix) {
            for (int m = 0; m < 3; m++) {
                R am = R(0), ap = R(0);
                for (int w = 0; w < 3; w++) {
                    am = am + zmin(f.s[w]) * f.w[w][m];
                    ap = ap + zmax(f.s[w]) * f.w[w][m];
                }
                f.am[m] = am; f.ap[m] = ap;
            }
            return;
        }

        // entropy fix: amdq from the left-going parts
        do {
            R s0;
            if constexpr (NC > 0) s0 = ca[2 + mu] - ca[5];
            else s0 = MM::divp(qa[mu], qa[0]) - MM::sqrt_(g * qa[0]);
            if (s0 > R(0) && f.s[0] > R(0)) {
                f.am[0] = R(0); f.am[1] = R(0); f.am[2] = R(0);
                break;
            }
            R h1 = qa[0] + f.w[0][0];
            R hu1 = qa[mu] + f.w[0][mu];
            R s1 = MM::divp(hu1, h1) - MM::sqrt_(g * h1);
            R sfract;
            if (s0 < R(0) && s1 > R(0))
                sfract = s0 * MM::div(s1 - f.s[0], s1 - s0);
            else if (f.s[0] < R(0))
                sfract = f.s[0];
            else
                sfract = R(0);
            for (int m = 0; m < 3; m++) f.am[m] = sfract * f.w[0][m];
            if (f.s[1] > R(0)) break;
            for (int m = 0; m < 3; m++) f.am[m] = f.am[m] + f.s[1] * f.w[1][m];
            R s03;
            if constexpr (NC > 0) s03 = cb[2 + mu] + cb[5];
            else s03 = MM::divp(qb[mu], qb[0]) + MM::sqrt_(g * qb[0]);
            R h3 = qb[0] - f.w[2][0];
            R hu3 = qb[mu] - f.w[2][mu];
            R s3 = MM::divp(hu3, h3) + MM::sqrt_(g * h3);
            if (s3 < R(0) && s03 > R(0))
                sfract = s3 * MM::div(s03 - f.s[2], s03 - s3);
            else if (f.s[2] < R(0))
                sfract = f.s[2];
            else
                break;
            for (int m = 0; m < 3; m++) f.am[m] = f.am[m] + sfract * f.w[2][m];
        } while (0);
        for (int m = 0; m < 3; m++) {
            R df = R(0);
            for (int w = 0; w < 3; w++) df = df + f.s[w] * f.w[w][m];
            f.ap[m] = df - f.am[m];
        }
    }

    template <int DIR>
    __device__ static inline void rpt2(int, const R* asdq, const Trans& roe, const PhysParams&,
                                       R* bm, R* bp)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        if constexpr (sizeof(R) == 4) {
            rpt2_f32<DIR>(asdq, roe, bm, bp);
            return;
        }
        using MM = M<R>;
        const R u = roe.u, v = roe.v, a = roe.a;
        R a1, a3;
        R a2 = asdq[mu] - u * asdq[0];
        if constexpr (C5) {
            const R ha = MM::divp(R(0.5), a);
            a1 = ha * ((v + a) * asdq[0] - asdq[mv]);
            a3 = ha * (-(v - a) * asdq[0] + asdq[mv]);
        } else {
            a1 = MM::divp((v + a) * asdq[0] - asdq[mv], R(2) * a);
            a3 = MM::divp(-(v - a) * asdq[0] + asdq[mv], R(2) * a);
        }
        R wb[3][3], sb[3];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a); sb[0] = v - a;
        wb[1][0] = R(0); wb[1][mu] = a2; wb[1][mv] = R(0); sb[1] = v;
        wb[2][0] = a3; wb[2][mu] = a3 * u; wb[2][mv] = a3 * (v + a); sb[2] = v + a;
        for (int m = 0; m < 3; m++) {
            R b0 = R(0), b1 = R(0);
            for (int w = 0; w < 3; w++) {
                b0 = b0 + zmin(sb[w]) * wb[w][m];
                b1 = b1 + zmax(sb[w]) * wb[w][m];
            }
            bm[m] = b0; bp[m] = b1;
        }
    }
};

// ---------------------------------------------------------------------------
// Euler equations + advected tracer, Roe solver with entropy fix.
// ---------------------------------------------------------------------------
template <class R, bool C5> struct Euler5T {
    static constexpr int ME = 5, MW = 5, MAUX = 0;
    static constexpr bool HAS_CONS = false;
    static constexpr bool FW = false;  // f-wave edge-velocity form (FwCons)
    using Face = FaceT<R, 5, 5>;
    struct Trans { R u, v, enth, a, g1a2, euv, u2v2; };

    static constexpr int NC = 0;  // no per-cell data
    __device__ static inline void cell(const R*, const PhysParams&, R*) {}
    // component m of wave w is zero by construction: the acoustic and
    // entropy waves carry no tracer, the shear wave only the transverse
    // momentum and energy, the tracer wave only the tracer
    template <int DIR> __device__ static constexpr bool wz(int w, int m)
    {
        return w == 4 ? m != 4 : (m == 4 || (w == 1 && (m == 0 || m == (DIR == 0 ? 1 : 2))));
    }
    static constexpr bool RPT2_SCALED = false;
    static constexpr bool SMAX = false;
    __device__ static inline R smax(const Face&, const Trans&) { return R(0); }
    template <int DIR> __device__ static void rpt2_scaled(const R*, const Trans&, R, R*, R*) {}

    template <int DIR>
    __device__ static inline void rpn2(const R* qa, const R* qb, const R*, const R*, R,
                                       const PhysParams& pp, Face& f, Trans& t)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        using MM = M<R>;
        const R gam = R(pp.gamma);
        const R g1 = R(pp.gamma - 1.0);
        R rl = MM::sqrt_(qa[0]);
        R rr = MM::sqrt_(qb[0]);
        R pl = g1 * (qa[3] - MM::divp(R(0.5) * (qa[1] * qa[1] + qa[2] * qa[2]), qa[0]));
        R pr = g1 * (qb[3] - MM::divp(R(0.5) * (qb[1] * qb[1] + qb[2] * qb[2]), qb[0]));
        R rsq2 = rl + rr;
        R u = MM::divp(MM::divp(qa[mu], rl) + MM::divp(qb[mu], rr), rsq2);
        R v = MM::divp(MM::divp(qa[mv], rl) + MM::divp(qb[mv], rr), rsq2);
        R enth = MM::divp(MM::divp(qa[3] + pl, rl) + MM::divp(qb[3] + pr, rr), rsq2);
        R u2v2 = u * u + v * v;
        R a2c = g1 * (enth - R(.5) * u2v2);
        R a = MM::sqrt_(a2c);
        R g1a2 = MM::divp(g1, a2c);
        R euv = enth - u2v2;
        t.u = u; t.v = v; t.enth = enth; t.a = a; t.euv = euv; t.u2v2 = u2v2;
        if constexpr (C5) t.g1a2 = MM::divp(g1, a * a);  // transverse split only
        else t.g1a2 = g1a2;

        R d0 = qb[0] - qa[0], d1 = qb[mu] - qa[mu], d2 = qb[mv] - qa[mv], d3 = qb[3] - qa[3];
        R a3 = g1a2 * (euv * d0 + u * d1 + v * d2 - d3);
        R a2 = d2 - v * d0;
        R a4 = MM::divp(d1 + (a - u) * d0 - a * a3, R(2) * a);
        R a1 = d0 - a3 - a4;

        f.w[0][0] = a1; f.w[0][mu] = a1 * (u - a); f.w[0][mv] = a1 * v;
        f.w[0][3] = a1 * (enth - u * a); f.w[0][4] = R(0); f.s[0] = u - a;
        f.w[1][0] = R(0); f.w[1][mu] = R(0); f.w[1][mv] = a2;
        f.w[1][3] = a2 * v; f.w[1][4] = R(0); f.s[1] = u;
        f.w[2][0] = a3; f.w[2][mu] = a3 * u; f.w[2][mv] = a3 * v;
        f.w[2][3] = a3 * R(0.5) * u2v2; f.w[2][4] = R(0); f.s[2] = u;
        f.w[3][0] = a4; f.w[3][mu] = a4 * (u + a); f.w[3][mv] = a4 * v;
        f.w[3][3] = a4 * (enth + u * a); f.w[3][4] = R(0); f.s[3] = u + a;
        f.w[4][0] = R(0); f.w[4][mu] = R(0); f.w[4][mv] = R(0);
        f.w[4][3] = R(0); f.w[4][4] = qb[4] - qa[4]; f.s[4] = u;

        if (!pp.efix) {
            for (int m = 0; m < 5; m++) {
                R am = R(0), ap = R(0);
                for (int w = 0; w < 5; w++) {
                    if (f.s[w] < R(0))
                        am = am + f.s[w] * f.w[w][m];
                    else
                        ap = ap + f.s[w] * f.w[w][m];
                }
                f.am[m] = am; f.ap[m] = ap;
            }
            return;
        }

        do {
            R rhoim1 = qa[0];
            R pim1 = g1 * (qa[3] - MM::divp(R(0.5) * (qa[mu] * qa[mu] + qa[mv] * qa[mv]), rhoim1));
            R cim1 = MM::sqrt_(MM::divp(gam * pim1, rhoim1));
            R s0 = MM::divp(qa[mu], rhoim1) - cim1;
            if (s0 >= R(0) && f.s[0] > R(0)) {
                for (int m = 0; m < 5; m++) f.am[m] = R(0);
                break;
            }
            R rho1 = qa[0] + f.w[0][0];
            R rhou1 = qa[mu] + f.w[0][mu];
            R rhov1 = qa[mv] + f.w[0][mv];
            R en1 = qa[3] + f.w[0][3];
            R p1 = g1 * (en1 - MM::divp(R(0.5) * (rhou1 * rhou1 + rhov1 * rhov1), rho1));
            R c1 = MM::sqrt_(MM::divp(gam * p1, rho1));
            R s1 = MM::divp(rhou1, rho1) - c1;
            R sfract;
            if (s0 < R(0) && s1 > R(0))
                sfract = MM::div(s0 * (s1 - f.s[0]), s1 - s0);
            else if (f.s[0] < R(0))
                sfract = f.s[0];
            else
                sfract = R(0);
            for (int m = 0; m < 5; m++) f.am[m] = sfract * f.w[0][m];
            if (f.s[1] >= R(0)) break;
            for (int m = 0; m < 5; m++) {
                f.am[m] = f.am[m] + f.s[1] * f.w[1][m];
                f.am[m] = f.am[m] + f.s[2] * f.w[2][m];
                f.am[m] = f.am[m] + f.s[4] * f.w[4][m];
            }
            R rhoi = qb[0];
            R pi_ = g1 * (qb[3] - MM::divp(R(0.5) * (qb[mu] * qb[mu] + qb[mv] * qb[mv]), rhoi));
            R ci = MM::sqrt_(MM::divp(gam * pi_, rhoi));
            R s3 = MM::divp(qb[mu], rhoi) + ci;
            R rho2 = qb[0] - f.w[3][0];
            R rhou2 = qb[mu] - f.w[3][mu];
            R rhov2 = qb[mv] - f.w[3][mv];
            R en2 = qb[3] - f.w[3][3];
            R p2 = g1 * (en2 - MM::divp(R(0.5) * (rhou2 * rhou2 + rhov2 * rhov2), rho2));
            R c2 = MM::sqrt_(MM::divp(gam * p2, rho2));
            R s2 = MM::divp(rhou2, rho2) + c2;
            if (s2 < R(0) && s3 > R(0))
                sfract = MM::div(s2 * (s3 - f.s[3]), s3 - s2);
            else if (f.s[3] < R(0))
                sfract = f.s[3];
            else
                break;
            for (int m = 0; m < 5; m++) f.am[m] = f.am[m] + sfract * f.w[3][m];
        } while (0);
        for (int m = 0; m < 5; m++) {
            R df = R(0);
            for (int w = 0; w < 5; w++) df = df + f.s[w] * f.w[w][m];
            f.ap[m] = df - f.am[m];
        }
    }

    template <int DIR>
    __device__ static inline void rpt2(int, const R* asdq, const Trans& t, const PhysParams&,
                                       R* bm, R* bp)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        using MM = M<R>;
        const R u = t.u, v = t.v, enth = t.enth, a = t.a, g1a2 = t.g1a2, euv = t.euv,
                u2v2 = t.u2v2;
        R a3 = g1a2 * (euv * asdq[0] + u * asdq[mu] + v * asdq[mv] - asdq[3]);
        R a2 = asdq[mu] - u * asdq[0];
        R a4 = MM::divp(asdq[mv] + (a - v) * asdq[0] - a * a3, R(2) * a);
        R a1 = asdq[0] - a3 - a4;
        R wb[4][5], sb[4];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a);
        wb[0][3] = a1 * (enth - v * a); wb[0][4] = R(0); sb[0] = v - a;
        wb[1][0] = a3; wb[1][mu] = a3 * u + a2; wb[1][mv] = a3 * v;
        wb[1][3] = a3 * R(0.5) * u2v2 + a2 * u; wb[1][4] = R(0); sb[1] = v;
        wb[2][0] = a4; wb[2][mu] = a4 * u; wb[2][mv] = a4 * (v + a);
        wb[2][3] = a4 * (enth + v * a); wb[2][4] = R(0); sb[2] = v + a;
        wb[3][0] = R(0); wb[3][mu] = R(0); wb[3][mv] = R(0); wb[3][3] = R(0);
        wb[3][4] = asdq[4]; sb[3] = v;
        for (int m = 0; m < 5; m++) {
            R b0 = R(0), b1 = R(0);
            for (int w = 0; w < 4; w++) {
                if (sb[w] < R(0))
                    b0 = b0 + sb[w] * wb[w][m];
                else
                    b1 = b1 + sb[w] * wb[w][m];
            }
            bm[m] = b0; bp[m] = b1;
        }
    }
};

// ---------------------------------------------------------------------------
// Euler equations + conserved tracer, HLLE: two waves with Einfeldt's speed
// bounds (Harten, Lax and van Leer 1983; Einfeldt 1988; LeVeque, Finite
// Volume Methods for Hyperbolic Problems, section 15.3.7).  The reference has
// no HLLE (applications/clawpack/euler/2d/rp/README); this solver takes the
// interface of clawpack46_rpn2_euler5.f / clawpack46_rpt2_euler5.f; its
// transverse split is the HLL form of rpt2_euler5's Roe split (rpt2 below).
//   s1 = min(u_a - c_a, u_roe - a_roe),  s2 = max(u_b + c_b, u_roe + a_roe)
//   W1 = (s2 (q_b - q_a) - (f(q_b) - f(q_a))) / (s2 - s1),  W2 = (q_b - q_a) - W1
// Equal states give exactly zero waves; s1 W1 + s2 W2 = f(q_b) - f(q_a).
// Field 5 is a conserved tracer density with flux u q_5.
// ---------------------------------------------------------------------------
template <class R> struct EulerHlle {
    static constexpr int ME = 5, MW = 2, MAUX = 0;
    static constexpr bool HAS_CONS = false;
    static constexpr bool FW = false;  // f-wave edge-velocity form (FwCons)
    using Face = FaceT<R, 5, 2>;
    using Trans = typename Euler5T<R, false>::Trans;

    static constexpr int NC = 0;  // no per-cell data
    __device__ static inline void cell(const R*, const PhysParams&, R*) {}
    template <int DIR> __device__ static constexpr bool wz(int, int) { return false; }
    static constexpr bool RPT2_SCALED = false;
    static constexpr bool SMAX = false;
    __device__ static inline R smax(const Face&, const Trans&) { return R(0); }
    template <int DIR> __device__ static void rpt2_scaled(const R*, const Trans&, R, R*, R*) {}

    // normal flux, normal velocity and sound speed of one state
    template <int DIR>
    __device__ static inline void flux(const R* q, R g1, R gam, R* f, R& un, R& c)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        using MM = M<R>;
        const R u = MM::divp(q[mu], q[0]);
        const R p = g1 * (q[3] - MM::divp(R(0.5) * (q[1] * q[1] + q[2] * q[2]), q[0]));
        f[0] = q[mu];
        f[mu] = q[mu] * u + p;
        f[mv] = q[mv] * u;
        f[3] = u * (q[3] + p);
        f[4] = q[4] * u;
        un = u;
        c = MM::sqrt_(MM::divp(gam * p, q[0]));
    }

    template <int DIR>
    __device__ static inline void rpn2(const R* qa, const R* qb, const R*, const R*, R,
                                       const PhysParams& pp, Face& f, Trans& t)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        using MM = M<R>;
        const R gam = R(pp.gamma);
        const R g1 = R(pp.gamma - 1.0);
        // Roe averages (the operations of Euler5T::rpn2)
        R rl = MM::sqrt_(qa[0]);
        R rr = MM::sqrt_(qb[0]);
        R pl = g1 * (qa[3] - MM::divp(R(0.5) * (qa[1] * qa[1] + qa[2] * qa[2]), qa[0]));
        R pr = g1 * (qb[3] - MM::divp(R(0.5) * (qb[1] * qb[1] + qb[2] * qb[2]), qb[0]));
        R rsq2 = rl + rr;
        R u = MM::divp(MM::divp(qa[mu], rl) + MM::divp(qb[mu], rr), rsq2);
        R v = MM::divp(MM::divp(qa[mv], rl) + MM::divp(qb[mv], rr), rsq2);
        R enth = MM::divp(MM::divp(qa[3] + pl, rl) + MM::divp(qb[3] + pr, rr), rsq2);
        R u2v2 = u * u + v * v;
        R a2c = g1 * (enth - R(.5) * u2v2);
        R a = MM::sqrt_(a2c);
        t.u = u; t.v = v; t.enth = enth; t.a = a; t.euv = enth - u2v2; t.u2v2 = u2v2;
        t.g1a2 = MM::divp(g1, a2c);

        R fa[5], fb[5], ua, ca, ub, cb;
        flux<DIR>(qa, g1, gam, fa, ua, ca);
        flux<DIR>(qb, g1, gam, fb, ub, cb);
        const R s1 = dmin(ua - ca, u - a);
        const R s2 = dmax(ub + cb, u + a);
        f.s[0] = s1;
        f.s[1] = s2;
        for (int m = 0; m < 5; m++) {
            const R dq = qb[m] - qa[m];
            const R w1 = MM::divp(s2 * dq - (fb[m] - fa[m]), s2 - s1);
            const R w2 = dq - w1;
            f.w[0][m] = w1;
            f.w[1][m] = w2;
            f.am[m] = zmin(s1) * w1 + zmin(s2) * w2;
            f.ap[m] = zmax(s1) * w1 + zmax(s2) * w2;
        }
    }

    // Transverse split: the HLL form of the Roe transverse matrix B at the
    // Roe average.  B asdq is the Roe split summed over its waves; with the
    // bounds t1 = v - a, t2 = v + a,
    //   V1 = (t2 asdq - B asdq) / (t2 - t1),  V2 = asdq - V1,
    //   bm = min(t1, 0) V1 + min(t2, 0) V2,  bp = max(t1, 0) V1 + max(t2, 0) V2,
    // so bm + bp = B asdq, and a fluctuation the Roe split gives no transverse
    // speed (a contact's) still spreads with the acoustic speeds, as the
    // normal solver spreads it (with the Roe split alone the scheme is
    // unstable at Courant numbers above about 0.5).
    template <int DIR>
    __device__ static inline void rpt2(int, const R* asdq, const Trans& t, const PhysParams&, R* bm,
                                       R* bp)
    {
        constexpr int mu = DIR == 0 ? 1 : 2;
        constexpr int mv = DIR == 0 ? 2 : 1;
        using MM = M<R>;
        const R u = t.u, v = t.v, enth = t.enth, a = t.a, g1a2 = t.g1a2, euv = t.euv, u2v2 = t.u2v2;
        R a3 = g1a2 * (euv * asdq[0] + u * asdq[mu] + v * asdq[mv] - asdq[3]);
        R a2 = asdq[mu] - u * asdq[0];
        R a4 = MM::divp(asdq[mv] + (a - v) * asdq[0] - a * a3, R(2) * a);
        R a1 = asdq[0] - a3 - a4;
        R wb[4][5], sb[4];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a);
        wb[0][3] = a1 * (enth - v * a); wb[0][4] = R(0); sb[0] = v - a;
        wb[1][0] = a3; wb[1][mu] = a3 * u + a2; wb[1][mv] = a3 * v;
        wb[1][3] = a3 * R(0.5) * u2v2 + a2 * u; wb[1][4] = R(0); sb[1] = v;
        wb[2][0] = a4; wb[2][mu] = a4 * u; wb[2][mv] = a4 * (v + a);
        wb[2][3] = a4 * (enth + v * a); wb[2][4] = R(0); sb[2] = v + a;
        wb[3][0] = R(0); wb[3][mu] = R(0); wb[3][mv] = R(0); wb[3][3] = R(0);
        wb[3][4] = asdq[4]; sb[3] = v;
        const R t1 = v - a, t2 = v + a;
        for (int m = 0; m < 5; m++) {
            R bq = R(0);
            for (int w = 0; w < 4; w++) bq = bq + sb[w] * wb[w][m];
            const R d = asdq[m];
            const R v1 = MM::divp(t2 * d - bq, t2 - t1);
            const R v2 = d - v1;
            bm[m] = zmin(t1) * v1 + zmin(t2) * v2;
            bp[m] = zmax(t1) * v1 + zmax(t2) * v2;
        }
    }
};

template <class R> using Swe = SweT<R, false>;
template <class R> using Swe5 = SweT<R, true>;
template <class R> using Euler5 = Euler5T<R, false>;
template <class R> using Euler5C5 = Euler5T<R, true>;

// Limiter phi(r) of the reference's inline limiter
// (src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_inlinelimiter.f:66-108).
template <class R> __device__ inline R limiter_phi(int mthlim, R r)
{
    switch (mthlim) {
    case 1: return zmax0(dmin(R(1), r));
    case 2: return dmax(zmax0(dmin(R(1), R(2) * r)), dmin(R(2), r));
    case 3: return M<R>::divp(r + M<R>::abs_(r), R(1) + M<R>::abs_(r));
    case 4: {
        R c = (R(1) + r) / R(2);
        return zmax0(dmin(dmin(c, R(2)), R(2) * r));
    }
    case 5: return r;
    case 6: {
        const R th = R((double)1.3f);  // single-precision literal in the reference
        return zmax0(dmin(dmin(th * r, (R(1) + r) / R(2)), th));
    }
    default: return R(1);
    }
}

}  // namespace hc
