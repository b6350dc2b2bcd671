/*
 * bem_detmath.h -- the arithmetic contract of the BEM kernels.
 *
 * Every floating-point operation that decides an ACA pivot must round the
 * same way on the host (the CPU restatement used as test oracle, host-side
 * set-up code) and on the gfx950 device.  On the symmetric meshes this code
 * is used with (icospheres, ellipsoids) many matrix entries are equal in
 * exact arithmetic, so a one-ulp difference between a host libm and the
 * device math library flips argmax ties and changes the ACA factors at the
 * ACA-tolerance level.  This header therefore fixes:
 *
 *   - complex arithmetic with explicit fma placement (both sides compile with
 *     -ffp-contract=off, so no other fusion happens);
 *   - det_sincos(): Cody-Waite reduction by pi/2 (three-part constant, fma)
 *     followed by Taylor polynomials on [-pi/4, pi/4] (terms to y^17 / y^18,
 *     truncation error < 1e-19), evaluated with Horner/fma;
 *   - bem_point_accumulate(): one quadrature point of the four Helmholtz
 *     collocation operators (single layer L, double layer M, adjoint double
 *     layer Mt, hypersingular N) as written in the reference's regular branch
 *     (boundaryelementsolver.cpp:867-888) and near-singular branch
 *     (boundaryelementsolver.cpp:1486-1506):
 *         G   = e^{ikr} / (4 pi r)
 *         G'  = (G / r)(ikr - 1)
 *         G'' = (G / r^2)(2 - 2ikr - k^2 r^2)
 *         L  += w G,  M += w G' rnq,  Mt += w G' rnp,
 *         N  += w [G' (-(np.nq + rnp rnq)/r) + G'' rnp rnq]
 *     with rnq = -(R.nq)/r, rnp = (R.np)/r, R = p - q.  Divisions by r are
 *     replaced by one reciprocal (round-off level reordering).
 *
 * sqrt and division are IEEE correctly rounded on both x86-64 (SSE2) and
 * gfx950 (HIP's default f64 lowering), and fma is exact-then-round on both.
 * The functions below contain nothing else, so the host and the device
 * produce bit-identical results.
 */
#ifndef BEM_DETMATH_H
#define BEM_DETMATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIP__)
#define BEM_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define BEM_HD static inline
#endif

typedef struct { double re, im; } bem_c64;

BEM_HD bem_c64 c64(double re, double im) { bem_c64 z; z.re = re; z.im = im; return z; }
BEM_HD bem_c64 c64_add(bem_c64 a, bem_c64 b) { return c64(a.re + b.re, a.im + b.im); }
BEM_HD bem_c64 c64_sub(bem_c64 a, bem_c64 b) { return c64(a.re - b.re, a.im - b.im); }
BEM_HD bem_c64 c64_neg(bem_c64 a) { return c64(-a.re, -a.im); }
BEM_HD bem_c64 c64_conj(bem_c64 a) { return c64(a.re, -a.im); }
BEM_HD bem_c64 c64_scale(bem_c64 a, double s) { return c64(a.re * s, a.im * s); }
/* (a.re + i a.im)(b.re + i b.im) with one fma per component */
BEM_HD bem_c64 c64_mul(bem_c64 a, bem_c64 b)
{
    return c64(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
}
/* acc + a*b */
BEM_HD bem_c64 c64_fma(bem_c64 a, bem_c64 b, bem_c64 acc)
{
    return c64(fma(a.re, b.re, fma(-a.im, b.im, acc.re)), fma(a.re, b.im, fma(a.im, b.re, acc.im)));
}
/* acc - a*b */
BEM_HD bem_c64 c64_fms(bem_c64 a, bem_c64 b, bem_c64 acc)
{
    return c64(fma(-a.re, b.re, fma(a.im, b.im, acc.re)), fma(-a.re, b.im, fma(-a.im, b.re, acc.im)));
}
/* acc + w*a (w real) */
BEM_HD bem_c64 c64_axpy(double w, bem_c64 a, bem_c64 acc) { return c64(fma(w, a.re, acc.re), fma(w, a.im, acc.im)); }
/* |a|^2, the quantity every ACA argmax compares */
BEM_HD double c64_abs2(bem_c64 a) { return fma(a.re, a.re, a.im * a.im); }
/* a / b via a * conj(b) / |b|^2 (operands are boundary-condition
   coefficients of moderate magnitude) */
BEM_HD bem_c64 c64_div(bem_c64 a, bem_c64 b)
{
    double d = c64_abs2(b);
    return c64(fma(a.re, b.re, a.im * b.im) / d, fma(a.im, b.re, -(a.re * b.im)) / d);
}

#define BEM_PI        3.141592653589793
#define BEM_INV_4PI   0.07957747154594767
#define BEM_INV_2PI   0.15915494309189535
#define BEM_2_OVER_PI 0.6366197723675814
#define BEM_PIO2_1    1.5707963267948966
#define BEM_PIO2_2    6.123233995736766e-17
#define BEM_PIO2_3   -1.4973849048591698e-33

/* sin and cos of x (|x| < 1e6).  Max error about one ulp (checked against
   extended-precision sin/cos by the test suite). */
BEM_HD void det_sincos(double x, double* s_out, double* c_out)
{
    double n = rint(x * BEM_2_OVER_PI);
    double y = fma(-n, BEM_PIO2_1, x);
    y = fma(-n, BEM_PIO2_2, y);
    y = fma(-n, BEM_PIO2_3, y);
    double z = y * y;
    /* sin(y) = y + y z (-1/3! + z/5! - ... + z^7/17!) */
    double ps = 2.8114572543455206e-15;                /*  1/17! */
    ps = fma(ps, z, -7.647163731819816e-13);           /* -1/15! */
    ps = fma(ps, z, 1.6059043836821613e-10);           /*  1/13! */
    ps = fma(ps, z, -2.505210838544172e-08);           /* -1/11! */
    ps = fma(ps, z, 2.7557319223985893e-06);           /*  1/9!  */
    ps = fma(ps, z, -0.0001984126984126984);           /* -1/7!  */
    ps = fma(ps, z, 0.008333333333333333);             /*  1/5!  */
    ps = fma(ps, z, -0.16666666666666666);             /* -1/3!  */
    double s = fma(y * z, ps, y);
    /* cos(y) = 1 - z/2 + z^2 (1/4! - z/6! + ... + z^7/18!) */
    double pc = -1.5619206968586225e-16;               /* -1/18! */
    pc = fma(pc, z, 4.779477332387385e-14);            /*  1/16! */
    pc = fma(pc, z, -1.1470745597729725e-11);          /* -1/14! */
    pc = fma(pc, z, 2.08767569878681e-09);             /*  1/12! */
    pc = fma(pc, z, -2.755731922398589e-07);           /* -1/10! */
    pc = fma(pc, z, 2.48015873015873e-05);             /*  1/8!  */
    pc = fma(pc, z, -0.001388888888888889);            /* -1/6!  */
    pc = fma(pc, z, 0.041666666666666664);             /*  1/4!  */
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    double c = w + (((1.0 - w) - hz) + (z * z) * pc);
    /* quadrant q = n mod 4 (|n| < 2^31 for |x| < 1e6):
       (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s) for q = 0..3.
       Selected without branches (odd q swaps, q >= 2 flips both sign
       bits), so that the device evaluates neighbouring points' chains
       interleaved; negation is a sign-bit flip, so the values are those of
       the four-way case split. */
    int q = (int)n;
    union { double d; uint64_t u; } so, co;
    so.d = (q & 1) ? c : s;
    co.d = (q & 1) ? -s : c;
    uint64_t flip = (uint64_t)(q & 2) << 62;
    so.u ^= flip;
    co.u ^= flip;
    *s_out = so.d;
    *c_out = co.d;
}

/* e^{i x} */
BEM_HD bem_c64 det_expi(double x)
{
    double s, c;
    det_sincos(x, &s, &c);
    return c64(c, s);
}

/* Accumulators of the four collocation operators of one (row, column) pair. */
typedef struct { bem_c64 L, M, Mt, N; } bem_ops4;

BEM_HD bem_ops4 bem_ops4_zero(void)
{
    bem_ops4 o;
    o.L = c64(0.0, 0.0); o.M = o.L; o.Mt = o.L; o.N = o.L;
    return o;
}

/* One quadrature point q with weight w for target p (normal np) and source
   normal nq; upnq = np.nq.  k is the (real) wavenumber. */
BEM_HD void bem_point_accumulate(bem_ops4* acc, double k,
                                 double px, double py, double pz,
                                 double npx, double npy, double npz,
                                 double nqx, double nqy, double nqz, double upnq,
                                 double qx, double qy, double qz, double w)
{
    double Rx = px - qx, Ry = py - qy, Rz = pz - qz;
    double rr = fma(Rz, Rz, fma(Ry, Ry, Rx * Rx));
    double r = sqrt(rr);
    double inv_r = 1.0 / r;
    double rnq = -fma(Rz, nqz, fma(Ry, nqy, Rx * nqx)) * inv_r;
    double rnp = fma(Rz, npz, fma(Ry, npy, Rx * npx)) * inv_r;
    double kr = k * r;
    double s, c;
    det_sincos(kr, &s, &c);
    double g = inv_r * BEM_INV_4PI;
    double Gre = c * g, Gim = s * g;
    double Grre = Gre * inv_r, Grim = Gim * inv_r;           /* G/r */
    double dGre = fma(-Grim, kr, -Grre);                     /* (G/r)(ikr-1) */
    double dGim = fma(Grre, kr, -Grim);
    double Grrre = Grre * inv_r, Grrim = Grim * inv_r;       /* G/r^2 */
    double a = fma(-kr, kr, 2.0);                            /* 2 - k^2 r^2 */
    double b = -2.0 * kr;                                    /* -2kr */
    double d2Gre = fma(Grrre, a, -(Grrim * b));
    double d2Gim = fma(Grrre, b, Grrim * a);
    double pp = rnp * rnq;
    double t = -(upnq + pp) * inv_r;
    double Nre = fma(dGre, t, d2Gre * pp);
    double Nim = fma(dGim, t, d2Gim * pp);
    acc->L.re = fma(w, Gre, acc->L.re);
    acc->L.im = fma(w, Gim, acc->L.im);
    acc->M.re = fma(w, dGre * rnq, acc->M.re);
    acc->M.im = fma(w, dGim * rnq, acc->M.im);
    acc->Mt.re = fma(w, dGre * rnp, acc->Mt.re);
    acc->Mt.im = fma(w, dGim * rnp, acc->Mt.im);
    acc->N.re = fma(w, Nre, acc->N.re);
    acc->N.im = fma(w, Nim, acc->N.im);
}

/* The order-5 7-point triangle rule (trianglequadraturerules.cpp:47-56):
   rows (xi1, xi2, w) with the reference's 14-digit literals; the point is
   (xi1 a + xi2 b) + xi3 c with xi3 = (1 - xi2) - xi1. */
#define BEM_TRI7 { \
    {0.33333333333333, 0.33333333333333, 0.22500000000000}, \
    {0.47014206410511, 0.47014206410511, 0.13239415278851}, \
    {0.47014206410511, 0.05971587178977, 0.13239415278851}, \
    {0.05971587178977, 0.47014206410511, 0.13239415278851}, \
    {0.10128650732346, 0.10128650732346, 0.12593918054483}, \
    {0.10128650732346, 0.79742698535309, 0.12593918054483}, \
    {0.79742698535309, 0.10128650732346, 0.12593918054483}}

/* One quadrature point of the field operators at an observation point x
   (BemOperatorField, boundaryelementsolver.cpp:2318-2353): with R = x - q,
   L += w G and M += w G' rnq, rnq = -(R.nq)/r; the same operation order as
   bem_point_accumulate. */
BEM_HD void bem_field_point_accumulate(bem_c64* L, bem_c64* M, double k, double px, double py, double pz,
                                       double nqx, double nqy, double nqz, double qx, double qy, double qz,
                                       double w)
{
    double Rx = px - qx, Ry = py - qy, Rz = pz - qz;
    double rr = fma(Rz, Rz, fma(Ry, Ry, Rx * Rx));
    double r = sqrt(rr);
    double inv_r = 1.0 / r;
    double rnq = -fma(Rz, nqz, fma(Ry, nqy, Rx * nqx)) * inv_r;
    double kr = k * r;
    double s, c;
    det_sincos(kr, &s, &c);
    double g = inv_r * BEM_INV_4PI;
    double Gre = c * g, Gim = s * g;
    double Grre = Gre * inv_r, Grim = Gim * inv_r;
    double dGre = fma(-Grim, kr, -Grre);
    double dGim = fma(Grre, kr, -Grim);
    L->re = fma(w, Gre, L->re);
    L->im = fma(w, Gim, L->im);
    M->re = fma(w, dGre * rnq, M->re);
    M->im = fma(w, dGim * rnq, M->im);
}

/* Contribution M phi_j - L dphi_j of element j to the field potential
   (calculateFieldSolutionRegular, boundaryelementsolver.cpp:2299-2304) */
BEM_HD bem_c64 bem_field_term(bem_c64 L, bem_c64 M, bem_c64 phi, bem_c64 dphi)
{
    return c64_sub(c64_mul(M, phi), c64_mul(L, dphi));
}

/* Boundary-condition substitution of one entry (reference
   implicitPhiMatrix2 / implicitDPhiMatrix2, boundaryelementsolver.cpp:2004-2040).
   X = L + cc Mt + delta cc/2,  Y = M + cc N - delta/2.
   Column coefficients (per column j, set once per frequency):
     subst_j (|alpha_j| < |beta_j|): Phi = cphi_j X,  DPhi = cd_j X - Y
       with cphi_j = -1/beta_j, cd_j = -alpha_j/beta_j
     otherwise:                     Phi = cphi_j Y,  DPhi = X + cd_j Y
       with cphi_j = 1/alpha_j,  cd_j = beta_j/alpha_j                    */
BEM_HD void bem_xy(const bem_ops4* o, bem_c64 cc, int diag, bem_c64* X, bem_c64* Y)
{
    bem_c64 x = c64_fma(cc, o->Mt, o->L);
    bem_c64 y = c64_fma(cc, o->N, o->M);
    if (diag) {
        x = c64_add(x, c64_scale(cc, 0.5));
        y.re = y.re - 0.5;
    }
    *X = x;
    *Y = y;
}

BEM_HD bem_c64 bem_phi_entry(bem_c64 X, bem_c64 Y, int subst, bem_c64 cphi)
{
    return subst ? c64_mul(cphi, X) : c64_mul(cphi, Y);
}

BEM_HD bem_c64 bem_dphi_entry(bem_c64 X, bem_c64 Y, int subst, bem_c64 cd)
{
    return subst ? c64_sub(c64_mul(cd, X), Y) : c64_fma(cd, Y, X);
}

/* Counter-based pivot generator: a splitmix64 finaliser over
   (seed, stream, block, draw).  Schedule independent and identical on host and
   device; replaces the reference's system-entropy QRandomGenerator draws
   (HMatrix/hmatrix.cpp:508, boundaryelementsolver.cpp:2156-2158). */
BEM_HD uint64_t bem_mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* uniform integer in [0, bound) for draw `draw` of `block` in `stream` */
BEM_HD int64_t bem_rand_bounded(uint64_t seed, uint64_t stream, uint64_t block, uint64_t draw, int64_t bound)
{
    uint64_t h = bem_mix64(seed ^ bem_mix64(stream * 0x100000001B3ull ^ bem_mix64(block ^ (draw << 40))));
    return (int64_t)(((h >> 32) * (uint64_t)bound) >> 32);
}

#endif /* BEM_DETMATH_H */
