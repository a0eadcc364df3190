// Host FP64 set-up for the device path (see geometry.h).  Built with the host compiler and
// -ffp-contract=off: every metric term is evaluated with the same operation sequence as the
// reference's Fortran expressions, so device results can be compared bit for bit with an
// FP64 CPU evaluation of the same algorithm.
#include "geometry.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// Panel frames of the reference's cube topology (topology/cubed_sphere_topology_mod.f90:29-55):
// index axes ex, ey, inner normal n and lattice origin r of panels 1..6.
const int EX[6][3] = {{-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 0}};
const int EY[6][3] = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
const int NV[6][3] = {{0, 0, 1}, {-1, 0, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
const int RV[6][3] = {{1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 0, 0}, {0, 0, 0}, {0, 1, 1}};

double pi_() { return std::acos(-1.0); }

// grid coordinate of Ah point i (1-based): alpha0 + (i-1)*hx, hx = (alpha1-alpha0)/N
double grid_hx(int N) { return (0.25 * pi_() - (-0.25 * pi_())) / (double)N; }
double grid_coord(int N, int i) { return -0.25 * pi_() + ((double)(i - 1) + 0.0) * grid_hx(N); }

// prototype-face vectors (ecs_metric_mod.f90:326-395)
void r_proto(double a, double b, double r[3]) {
    const double ta = std::tan(a), tb = std::tan(b);
    const double d = std::sqrt(1.0 + ta * ta + tb * tb);
    r[0] = ta / d;
    r[1] = tb / d;
    r[2] = 1.0 / d;
}
void a1_proto(double a, double b, double v[3]) {
    const double ta = std::tan(a), tb = std::tan(b);
    const double s = std::pow(1.0 + std::pow(ta, 2.0) + std::pow(tb, 2.0), 3.0 / 2.0);
    v[0] = (1.0 + std::pow(ta, 2.0)) * (1.0 + std::pow(tb, 2.0)) / s;
    v[1] = -ta * tb * (1.0 + std::pow(ta, 2.0)) / s;
    v[2] = -ta * (1.0 + std::pow(ta, 2.0)) / s;
}
void a2_proto(double a, double b, double v[3]) {
    const double ta = std::tan(a), tb = std::tan(b);
    const double s = std::pow(1.0 + std::pow(ta, 2.0) + std::pow(tb, 2.0), 3.0 / 2.0);
    v[0] = -ta * tb * (1.0 + std::pow(tb, 2.0)) / s;
    v[1] = (1.0 + std::pow(ta, 2.0)) * (1.0 + std::pow(tb, 2.0)) / s;
    v[2] = -tb * (1.0 + std::pow(tb, 2.0)) / s;
}
void b1_proto(double a, double b, double v[3]) {
    const double ta = std::tan(a), tb = std::tan(b);
    const double sg = std::sqrt(1.0 + ta * ta + tb * tb);
    v[0] = (1.0 + tb * tb - (ta * tb) * (ta * tb) / (1.0 + ta * ta)) / sg;
    v[1] = 0.0;
    v[2] = -ta * (1.0 + tb * tb / (1 + ta * ta)) / sg;
}
void b2_proto(double a, double b, double v[3]) {
    const double ta = std::tan(a), tb = std::tan(b);
    const double sg = std::sqrt(1.0 + ta * ta + tb * tb);
    v[0] = 0.0;
    v[1] = ((1 + ta * ta) - (ta * tb) * (ta * tb) / (1 + tb * tb)) / sg;
    v[2] = -tb * (ta * ta / (1 + tb * tb) + 1.0) / sg;
}
// prototype face -> panel p with the identity rotation (ecs_metric_mod.f90:303-324)
void to_panel(int p, const double r[3], double out[3]) {
    double t[3];
    for (int k = 0; k < 3; ++k)
        t[k] = (double)EX[p][k] * r[0] + (double)EY[p][k] * r[1] - (double)NV[p][k] * r[2];
    for (int c = 0; c < 3; ++c) {
        double s = (c == 0 ? 1.0 : 0.0) * t[0];
        s = s + (c == 1 ? 1.0 : 0.0) * t[1];
        s = s + (c == 2 ? 1.0 : 0.0) * t[2];
        out[c] = s;
    }
}
double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// integer lattice position of Ah point (i,j) (1-based, may lie outside 1..n) of panel p on
// the cube surface of edge length n-1; panels sharing a point give the same lattice position.
void lattice(int p, int n, int i, int j, int L[3]) {
    for (int k = 0; k < 3; ++k) L[k] = (i - 1) * EX[p][k] + (j - 1) * EY[p][k] + (n - 1) * RV[p][k];
}
int idot(const int a[3], const int b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

}  // namespace

extern "C" void swmh_geometry(int N, int pitch, double omega, const double axis[3], double *J, double *Qi11,
                              double *Qi12, double *Qi22, double *f, double *r, double *a1, double *a2) {
    const int n = N + 1;
    const long plane = 6L * n * pitch;
    for (int p = 0; p < 6; ++p)
        for (int j = 1; j <= n; ++j) {
            const double beta = grid_coord(N, j);
            for (int i = 1; i <= n; ++i) {
                const double alpha = grid_coord(N, i);
                const long id = ((long)p * n + (j - 1)) * pitch + (i - 1);
                const double ta = std::tan(alpha), tb = std::tan(beta);
                const double sg = std::sqrt(1.0 + ta * ta + tb * tb);
                if (J) J[id] = (1.0 + ta * ta) * (1.0 + tb * tb) / (sg * sg * sg);
                if (Qi11) Qi11[id] = sg * sg / (1.0 + ta * ta);
                if (Qi12) Qi12[id] = sg * sg * ta * tb / ((1 + ta * ta) * (1 + tb * tb));
                if (Qi22) Qi22[id] = sg * sg / (1.0 + tb * tb);
                double rp[3], rr[3], v[3], w[3];
                r_proto(alpha, beta, rp);
                to_panel(p, rp, rr);
                if (r)
                    for (int k = 0; k < 3; ++k) r[id + k * plane] = rr[k];
                if (f) {
                    const double pr = (axis[0] * rr[0] + axis[1] * rr[1] + axis[2] * rr[2]) /
                                      std::sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
                    f[id] = 2 * omega * pr;
                }
                if (a1) {
                    a1_proto(alpha, beta, v);
                    to_panel(p, v, w);
                    for (int k = 0; k < 3; ++k) a1[id + k * plane] = w[k];
                }
                if (a2) {
                    a2_proto(alpha, beta, v);
                    to_panel(p, v, w);
                    for (int k = 0; k < 3; ++k) a2[id + k * plane] = w[k];
                }
            }
        }
}

extern "C" void swmh_basis_b(int N, int pitch, double *b1, double *b2) {
    const int n = N + 1;
    const long plane = 6L * n * pitch;
    for (int p = 0; p < 6; ++p)
        for (int j = 1; j <= n; ++j) {
            const double beta = grid_coord(N, j);
            for (int i = 1; i <= n; ++i) {
                const double alpha = grid_coord(N, i);
                const long id = ((long)p * n + (j - 1)) * pitch + (i - 1);
                double v[3], w[3];
                b1_proto(alpha, beta, v);
                to_panel(p, v, w);
                for (int k = 0; k < 3; ++k) b1[id + k * plane] = w[k];
                b2_proto(alpha, beta, v);
                to_panel(p, v, w);
                for (int k = 0; k < 3; ++k) b2[id + k * plane] = w[k];
            }
        }
}

// create_latlon_interp (regrid_factory_mod.f90:183-301) on the Ah (xy) mesh: stencil bounds
// 1..n, hx as the mesh, shift 0; transform_cartesian_to_native_ecs (ecs_metric_mod.f90:480-506)
// for the panel and (alpha, beta); cubic Lagrange weights; for the rotation the covariant branch
// of create_latlon_vector_regrid (:159-176) with b1, b2 of calculate_ecs_b1_2d / _b2_2d.
extern "C" void swmh_latlon_table(int N, int Nlat, int Nlon, int *panel, int *x0, int *y0, double *wx, double *wy,
                                  double *rot) {
    const int n = N + 1;
    const double pi = pi_(), hx = grid_hx(N), alpha0 = -0.25 * pi;
    const double hlon = 2.0 * pi / (double)Nlon, hlat = pi / ((double)Nlat - 1.0);
    for (int j = 1; j <= Nlat; ++j) {
        const double lat = -0.5 * pi + hlat * (double)(j - 1);
        for (int i = 1; i <= Nlon; ++i) {
            const double lon = hlon * (double)(i - 1);
            const double r[3] = {std::cos(lat) * std::cos(lon), std::cos(lat) * std::sin(lon), std::sin(lat)};
            // panel with the largest projection on the outer normal (first one on ties)
            int pi_ind = 0;
            double dmax = 0.0;
            for (int p = 0; p < 6; ++p) {
                double s = r[0] * (double)NV[p][0];
                s = s + r[1] * (double)NV[p][1];
                s = s + r[2] * (double)NV[p][2];
                const double d = -s;
                if (d > dmax) {
                    pi_ind = p;
                    dmax = d;
                }
            }
            double sx = r[0] * (double)EX[pi_ind][0];
            sx = sx + r[1] * (double)EX[pi_ind][1];
            sx = sx + r[2] * (double)EX[pi_ind][2];
            double sy = r[0] * (double)EY[pi_ind][0];
            sy = sy + r[1] * (double)EY[pi_ind][1];
            sy = sy + r[2] * (double)EY[pi_ind][2];
            const double alpha = std::atan(sx / dmax), beta = std::atan(sy / dmax);
            const double zdx = (alpha - alpha0) / hx + 1.0 - 0.0;
            const double zdy = (beta - alpha0) / hx + 1.0 - 0.0;
            const int xi = std::max(std::min((int)std::floor(zdx) + 2, n) - 4 + 1, 1);
            const int yi = std::max(std::min((int)std::floor(zdy) + 2, n) - 4 + 1, 1);
            const long q = (long)(j - 1) * Nlon + (i - 1);
            panel[q] = pi_ind;
            x0[q] = xi - 1;
            y0[q] = yi - 1;
            const double zx = zdx - (double)xi, zy = zdy - (double)yi;
            const double z2[2] = {zx, zy};
            double *w2[2] = {wx + 4 * q, wy + 4 * q};
            for (int d = 0; d < 2; ++d) {
                const double z = z2[d];
                w2[d][0] = -((z - 1.0) * (z - 2.0) * (z - 3.0) / 6.0);
                w2[d][1] = z * (z - 2.0) * (z - 3.0) / 2.0;
                w2[d][2] = -(z * (z - 1.0) * (z - 3.0) / 2.0);
                w2[d][3] = z * (z - 1.0) * (z - 2.0) / 6.0;
            }
            if (rot) {
                const double iv[3] = {-std::sin(lon), std::cos(lon), 0.0};
                const double jv[3] = {-std::sin(lat) * std::cos(lon), -std::sin(lat) * std::sin(lon), std::cos(lat)};
                double v[3], b1[3], b2[3];
                b1_proto(alpha, beta, v);
                to_panel(pi_ind, v, b1);
                b2_proto(alpha, beta, v);
                to_panel(pi_ind, v, b2);
                rot[4 * q + 0] = dot3(iv, b1);
                rot[4 * q + 1] = dot3(jv, b1);
                rot[4 * q + 2] = dot3(iv, b2);
                rot[4 * q + 3] = dot3(jv, b2);
            }
        }
    }
}

extern "C" void swmh_sync_weights(int N, double *qb, double *qt, double *ql, double *qr) {
    const int n = N + 1;
    const double pi = pi_();
    const double hx = grid_hx(N);
    double a[3], b[3], t[3];
    for (int i = 1; i <= n; ++i) {
        const double al = -0.25 * pi + (i - 1) * hx;
        a2_proto(al, -0.25 * pi, a);
        b1_proto(al, 0.25 * pi, t);
        b[0] = t[0]; b[1] = -t[2]; b[2] = t[1];
        qb[i - 1] = dot3(a, b);
        a2_proto(al, 0.25 * pi, a);
        b1_proto(al, -0.25 * pi, t);
        b[0] = t[0]; b[1] = t[2]; b[2] = -t[1];
        qt[i - 1] = dot3(a, b);
    }
    for (int j = 1; j <= n; ++j) {
        const double be = -0.25 * pi + (j - 1) * hx;
        a1_proto(-0.25 * pi, be, a);
        b2_proto(0.25 * pi, be, t);
        b[0] = -t[2]; b[1] = t[1]; b[2] = t[0];
        ql[j - 1] = dot3(a, b);
        a1_proto(0.25 * pi, be, a);
        b2_proto(-0.25 * pi, be, t);
        b[0] = t[2]; b[1] = t[1]; b[2] = -t[0];
        qr[j - 1] = dot3(a, b);
    }
}

extern "C" int swmh_quad_weights(int N, int order, double *A) {
    const int n = N + 1;
    int ne = 0;
    double e[6] = {0, 0, 0, 0, 0, 0};
    if (order == 21) { ne = 2; e[0] = 0.5; e[1] = 1.0; }
    else if (order == 42) { ne = 4; e[0] = 17.0 / 48.0; e[1] = 59.0 / 48.0; e[2] = 43.0 / 48.0; e[3] = 49.0 / 48.0; }
    else if (order == 63) {
        ne = 6;
        e[0] = 0.13649e5 / 0.43200e5; e[1] = 0.12013e5 / 0.8640e4; e[2] = 0.2711e4 / 0.4320e4;
        e[3] = 0.5359e4 / 0.4320e4; e[4] = 0.7877e4 / 0.8640e4; e[5] = 0.43801e5 / 0.43200e5;
    } else {
        return -1;
    }
    for (int i = 0; i < n; ++i) A[i] = 1.0;
    for (int i = 0; i < ne && i < n; ++i) A[i] = e[i];
    for (int i = std::max(n - ne, 0); i < n; ++i) A[i] = e[n - 1 - i];
    return 0;
}

// Neighbour table derived from the cube lattice: a halo point beyond an edge of panel p lies,
// in the unfolded plane, d steps past the edge; on the cube it is the point d steps inside the
// adjacent face, i.e. lattice(edge point) + d * n_p (n_p = inner normal of p).  The adjacent
// face q across the i=1 edge has n_q = ex_p (i=n: -ex_p, j=1: ey_p, j=n: -ey_p).  Vector
// components follow the index directions: the local +i / +j directions at the halo point are
// lattice directions which coincide (up to sign) with q's ex or ey.
extern "C" void swmh_halo_table(int N, int ndepth, int *tab) {
    const int n = N + 1;
    for (int p = 0; p < 6; ++p)
        for (int e = 0; e < 4; ++e)
            for (int k = 0; k < n; ++k)
                for (int d = 0; d < ndepth; ++d) {
                    int L[3], nq[3], di[3], dj[3];
                    // edge point (1-based) and the lattice directions of local +i / +j there
                    int ie = 0, je = 0;
                    for (int c = 0; c < 3; ++c) {
                        switch (e) {
                        case 0: nq[c] = EX[p][c]; di[c] = -NV[p][c]; dj[c] = EY[p][c]; break;
                        case 1: nq[c] = -EX[p][c]; di[c] = NV[p][c]; dj[c] = EY[p][c]; break;
                        case 2: nq[c] = EY[p][c]; di[c] = EX[p][c]; dj[c] = -NV[p][c]; break;
                        default: nq[c] = -EY[p][c]; di[c] = EX[p][c]; dj[c] = NV[p][c]; break;
                        }
                    }
                    if (e == 0) { ie = 1; je = k + 1; }
                    if (e == 1) { ie = n; je = k + 1; }
                    if (e == 2) { ie = k + 1; je = 1; }
                    if (e == 3) { ie = k + 1; je = n; }
                    lattice(p, n, ie, je, L);
                    for (int c = 0; c < 3; ++c) L[c] += d * NV[p][c];
                    int q = -1;
                    for (int s = 0; s < 6; ++s)
                        if (NV[s][0] == nq[0] && NV[s][1] == nq[1] && NV[s][2] == nq[2]) q = s;
                    int rel[3];
                    for (int c = 0; c < 3; ++c) rel[c] = L[c] - (n - 1) * RV[q][c];
                    const int iq = idot(rel, EX[q]), jq = idot(rel, EY[q]);
                    // local axis +i maps to q's ex or ey
                    const int swap = idot(di, EX[q]) == 0 ? 1 : 0;
                    const int si = swap ? idot(di, EY[q]) : idot(di, EX[q]);
                    const int sj = swap ? idot(dj, EX[q]) : idot(dj, EY[q]);
                    int *t = tab + ((((long)p * 4 + e) * n + k) * ndepth + d) * 4;
                    t[0] = q;
                    t[1] = iq;
                    t[2] = jq;
                    t[3] = swap | (si < 0 ? 2 : 0) | (sj < 0 ? 4 : 0);
                }
}

// ---------------------------------------------------------------------------------------------
// Seeded spherical-harmonic initial state
namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    bool have = false;
    double spare = 0.0;
    double normal() {  // Box-Muller
        if (have) { have = false; return spare; }
        const double u1 = uniform(), u2 = uniform();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        spare = rad * std::sin(2.0 * pi_() * u2);
        have = true;
        return rad * std::cos(2.0 * pi_() * u2);
    }
};

// Fully normalised associated Legendre functions without the Condon-Shortley phase,
// Pb[l][m] at mu = sin(phi); Qb[l][m] = Pb[l][m]/cos(phi) for m >= 1 (regular at the poles).
struct Legendre {
    int L;
    std::vector<double> P, Q;  // (L+1)*(L+1)
    explicit Legendre(int L_) : L(L_), P((L_ + 1) * (L_ + 1)), Q((L_ + 1) * (L_ + 1)) {}
    double &p(int l, int m) { return P[l * (L + 1) + m]; }
    double &q(int l, int m) { return Q[l * (L + 1) + m]; }
    void eval(double mu, double c) {
        // m = 0 column
        p(0, 0) = 1.0;
        if (L >= 1) p(1, 0) = std::sqrt(3.0) * mu;
        for (int l = 2; l <= L; ++l) {
            const double a = std::sqrt((4.0 * l * l - 1.0) / (double)(l * l));
            const double b = std::sqrt(((l - 1.0) * (l - 1.0)) * (2.0 * l + 1.0) / ((2.0 * l - 3.0) * (double)(l * l)));
            p(l, 0) = a * mu * p(l - 1, 0) - b * p(l - 2, 0);
        }
        // m >= 1 columns on Q = P / cos(phi)
        double qmm = 0.0;
        for (int m = 1; m <= L; ++m) {
            qmm = (m == 1) ? std::sqrt(3.0) : std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * c * qmm;
            q(m, m) = qmm;
            if (m + 1 <= L) q(m + 1, m) = std::sqrt(2.0 * m + 3.0) * mu * qmm;
            for (int l = m + 2; l <= L; ++l) {
                const double a = std::sqrt((4.0 * l * l - 1.0) / (double)(l * l - m * m));
                const double b = std::sqrt(((l - 1.0) * (l - 1.0) - m * m) * (2.0 * l + 1.0) /
                                           ((2.0 * l - 3.0) * (double)(l * l - m * m)));
                q(l, m) = a * mu * q(l - 1, m) - b * q(l - 2, m);
            }
            for (int l = m; l <= L; ++l) p(l, m) = c * q(l, m);
        }
    }
};

}  // namespace

extern "C" int swmh_ic_sh(int N, int pitch, unsigned long long seed, int lmax, double h0, double h_amp,
                          double v_max, double *h, double *u, double *v, int nthreads) {
    if (N < 1 || lmax < 1 || pitch < N + 1) return -1;
    const int n = N + 1;
    const long plane = 6L * n * pitch;
    const int K = (lmax + 1) * (lmax + 1);
    std::vector<double> ch(2 * K, 0.0), cp(2 * K, 0.0);  // [l][m] cos, sin coefficients
    Rng rng{seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<double> &c = pass == 0 ? ch : cp;
        for (int l = 1; l <= lmax; ++l)
            for (int m = 0; m <= l; ++m) {
                const double amp = 1.0 / ((double)l * (double)l);
                c[2 * (l * (lmax + 1) + m)] = rng.normal() * amp;
                c[2 * (l * (lmax + 1) + m) + 1] = (m > 0) ? rng.normal() * amp : 0.0;
            }
    }
    std::vector<double> r(3 * plane), a1(3 * plane), a2(3 * plane);
    const double axis[3] = {0.0, 0.0, 1.0};
    swmh_geometry(N, pitch, 0.0, axis, nullptr, nullptr, nullptr, nullptr, nullptr, r.data(), a1.data(), a2.data());
    std::vector<double> hp(plane, 0.0), vx(plane, 0.0), vy(plane, 0.0), vz(plane, 0.0);
    auto work = [&](int t0, int tstride) {
        Legendre Lg(lmax);
        std::vector<double> cm(lmax + 1), sm(lmax + 1);
        for (long row = t0; row < 6L * n; row += tstride) {
            for (int i = 0; i < n; ++i) {
                const long id = row * pitch + i;
                const double x = r[id], y = r[id + plane], z = r[id + 2 * plane];
                const double rho = std::sqrt(x * x + y * y);
                const double nrm = std::sqrt(rho * rho + z * z);
                const double mu = z / nrm, c = rho / nrm;
                const double lam = std::atan2(y, x);
                Lg.eval(mu, c);
                for (int m = 0; m <= lmax; ++m) { cm[m] = std::cos(m * lam); sm[m] = std::sin(m * lam); }
                double hv = 0.0, dpsi_phi = 0.0, dpsi_lam_c = 0.0;
                for (int l = 1; l <= lmax; ++l) {
                    for (int m = 0; m <= l; ++m) {
                        const int kk = 2 * (l * (lmax + 1) + m);
                        const double cosm = cm[m], sinm = sm[m];
                        hv += Lg.p(l, m) * (ch[kk] * cosm + ch[kk + 1] * sinm);
                        const double ap = cp[kk], bp = cp[kk + 1];
                        double dphi;
                        if (m == 0) {
                            dphi = std::sqrt(l * (l + 1.0) / 2.0) * Lg.p(l, 1);
                        } else {
                            const double e = std::sqrt((2.0 * l + 1.0) * (double)(l * l - m * m) / (2.0 * l - 1.0));
                            dphi = -l * mu * Lg.q(l, m) + (l - 1 >= m ? e * Lg.q(l - 1, m) : 0.0);
                        }
                        dpsi_phi += dphi * (ap * cosm + bp * sinm);
                        if (m > 0) dpsi_lam_c += Lg.q(l, m) * m * (-ap * sinm + bp * cosm);
                    }
                }
                // V = r x grad psi = -dpsi/dphi e_lam + (1/cos phi) dpsi/dlam e_phi
                const double sl = std::sin(lam), cl = std::cos(lam);
                const double el[3] = {-sl, cl, 0.0};
                const double ep[3] = {-mu * cl, -mu * sl, c};
                hp[id] = hv;
                vx[id] = -dpsi_phi * el[0] + dpsi_lam_c * ep[0];
                vy[id] = -dpsi_phi * el[1] + dpsi_lam_c * ep[1];
                vz[id] = -dpsi_phi * el[2] + dpsi_lam_c * ep[2];
            }
        }
    };
    const int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t, nt);
    work(0, nt);
    for (auto &t : th) t.join();
    double hmax = 0.0, vmax = 0.0;
    for (long row = 0; row < 6L * n; ++row)
        for (int i = 0; i < n; ++i) {
            const long id = row * pitch + i;
            hmax = std::max(hmax, std::fabs(hp[id]));
            vmax = std::max(vmax, std::sqrt(vx[id] * vx[id] + vy[id] * vy[id] + vz[id] * vz[id]));
        }
    const double sh = hmax > 0 ? h_amp / hmax : 0.0, sv = vmax > 0 ? v_max / vmax : 0.0;
    for (long row = 0; row < 6L * n; ++row)
        for (int i = 0; i < n; ++i) {
            const long id = row * pitch + i;
            const double Vx = vx[id] * sv, Vy = vy[id] * sv, Vz = vz[id] * sv;
            if (h) h[id] = h0 + hp[id] * sh;
            // covariant projection sum([vx,vy,vz]*a) (test_fields_mod.f90:428)
            if (u) u[id] = Vx * a1[id] + Vy * a1[id + plane] + Vz * a1[id + 2 * plane];
            if (v) v[id] = Vx * a2[id] + Vy * a2[id + plane] + Vz * a2[id + 2 * plane];
        }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Edge syncs as explicit sparse matrices over the panel-edge points (used by the adjoint).
//
// Slot numbering of the 4n-4 edge points of a panel (0-based i, j):
//   j = 0: slot i;  j = n-1: slot n+i;  i = 0 (0<j<n-1): slot 2n+j-1;  i = n-1 (0<j<n-1): 3n+j-3.
// Global slot = p*(4n-4) + slot.  Scalar entries: (row slot, column slot, weight).  Vector
// entries index component c (0 = x, 1 = y) as c*6*(4n-4) + global slot.
//
// The scalar matrix is halo_Ah_scalar_sync_mod.f90:52-93 (halves at edges, thirds at cube
// corners) and the vector matrix the covariant ecs_Ah_vec_sync_mod.f90:77-124, called with
// (y, x) components as the gradient does; the neighbour's copy of a shared point and the
// component mapping of the vector halo come from the neighbour table above.
namespace {

int edge_slot(int n, int i, int j) {
    if (j == 0) return i;
    if (j == n - 1) return n + i;
    if (i == 0) return 2 * n + j - 1;
    return 3 * n + j - 3;
}

struct EdgeBuilder {
    int n, es;
    std::vector<int> tab;
    int *rows, *cols;
    double *w;
    int cap, nnz = 0;
    EdgeBuilder(int N, int *r, int *c, double *ww, int cp) : n(N + 1), es(4 * N), rows(r), cols(c), w(ww), cap(cp) {
        tab.assign((size_t)6 * 4 * n * 4, 0);
        swmh_halo_table(N, 1, tab.data());
    }
    const int *nb(int p, int e, int k) const { return &tab[(((size_t)p * 4 + e) * n + k) * 4]; }
    void add(int r, int c, double v) {
        if (nnz < cap) { rows[nnz] = r; cols[nnz] = c; w[nnz] = v; }
        ++nnz;
    }
    int gslot(int p, int i, int j) const { return p * es + edge_slot(n, i, j); }
    // neighbour copy of the shared point k of edge e: global slot
    int nslot(int p, int e, int k) const { const int *t = nb(p, e, k); return gslot(t[0], t[1], t[2]); }
    // component `c` (0 x, 1 y) of the vector halo at (e, k): neighbour component and sign
    void ncomp(int p, int e, int k, int c, int &col, double &sg) const {
        const int *t = nb(p, e, k);
        const int code = t[3];
        const int src = c == 0 ? ((code & 1) ? 1 : 0) : ((code & 1) ? 0 : 1);
        sg = c == 0 ? ((code & 2) ? -1.0 : 1.0) : ((code & 4) ? -1.0 : 1.0);
        col = src * 6 * es + gslot(t[0], t[1], t[2]);
    }
};

}  // namespace

extern "C" int swmh_edge_sync_matrix(int N, int vector, int *rows, int *cols, double *w, int cap) {
    EdgeBuilder b(N, rows, cols, w, cap);
    const int n = b.n, e = n - 1, V = 6 * b.es;
    if (!vector) {
        for (int p = 0; p < 6; ++p) {
            for (int m = 1; m < n - 1; ++m) {
                const int side[4][3] = {{0, m, 0}, {e, m, 1}, {m, 0, 2}, {m, e, 3}};  // i, j, edge
                for (auto &s : side) {
                    const int r = b.gslot(p, s[0], s[1]);
                    b.add(r, r, 0.5);
                    b.add(r, b.nslot(p, s[2], m), 0.5);
                }
            }
            const int corner[4][6] = {{0, 0, 0, 0, 2, 0}, {e, 0, 1, 0, 2, e}, {0, e, 0, e, 3, 0}, {e, e, 1, e, 3, e}};
            for (auto &c : corner) {
                const int r = b.gslot(p, c[0], c[1]);
                b.add(r, r, 1.0 / 3.0);
                b.add(r, b.nslot(p, c[2], c[3]), 1.0 / 3.0);
                b.add(r, b.nslot(p, c[4], c[5]), 1.0 / 3.0);
            }
        }
        return b.nnz;
    }
    std::vector<double> qb(n), qt(n), ql(n), qr(n);
    swmh_sync_weights(N, qb.data(), qt.data(), ql.data(), qr.data());
    auto X = [&](int s) { return s; };
    auto Y = [&](int s) { return V + s; };
    auto halo = [&](int row, int p, int ed, int k, int c, double wt) {
        int col;
        double sg;
        b.ncomp(p, ed, k, c, col, sg);
        b.add(row, col, sg * wt);
    };
    for (int p = 0; p < 6; ++p) {
        // cube corners: (i, j, first edge, k1, second edge, k2, weight arrays)
        {
            const int s = b.gslot(p, 0, 0);
            b.add(Y(s), Y(s), 1.0 / 3); halo(Y(s), p, 0, 0, 1, 1.0 / 3); halo(Y(s), p, 2, 0, 1, 1.0 / 3);
            halo(Y(s), p, 2, 0, 0, qb[0] / 3);
            b.add(X(s), X(s), 1.0 / 3); halo(X(s), p, 0, 0, 0, 1.0 / 3); halo(X(s), p, 2, 0, 0, 1.0 / 3);
            halo(X(s), p, 0, 0, 1, ql[0] / 3);
        }
        {
            const int s = b.gslot(p, e, 0);
            b.add(Y(s), Y(s), 1.0 / 3); halo(Y(s), p, 1, 0, 1, 1.0 / 3); halo(Y(s), p, 2, e, 1, 1.0 / 3);
            halo(Y(s), p, 2, e, 0, qb[e] / 3);
            b.add(X(s), X(s), 1.0 / 3); halo(X(s), p, 1, 0, 0, 1.0 / 3); halo(X(s), p, 2, e, 0, 1.0 / 3);
            halo(X(s), p, 1, 0, 1, qr[0] / 3);
        }
        {
            const int s = b.gslot(p, 0, e);
            b.add(Y(s), Y(s), 1.0 / 3); halo(Y(s), p, 0, e, 1, 1.0 / 3); halo(Y(s), p, 3, 0, 1, 1.0 / 3);
            halo(Y(s), p, 3, 0, 0, qt[0] / 3);
            b.add(X(s), X(s), 1.0 / 3); halo(X(s), p, 0, e, 0, 1.0 / 3); halo(X(s), p, 3, 0, 0, 1.0 / 3);
            halo(X(s), p, 0, e, 1, ql[e] / 3);
        }
        {
            const int s = b.gslot(p, e, e);
            b.add(Y(s), Y(s), 1.0 / 3); halo(Y(s), p, 1, e, 1, 1.0 / 3); halo(Y(s), p, 3, e, 1, 1.0 / 3);
            halo(Y(s), p, 3, e, 0, qt[e] / 3);
            b.add(X(s), X(s), 1.0 / 3); halo(X(s), p, 1, e, 0, 1.0 / 3); halo(X(s), p, 3, e, 0, 1.0 / 3);
            halo(X(s), p, 1, e, 1, qr[e] / 3);
        }
        for (int m = 1; m < n - 1; ++m) {
            int s = b.gslot(p, 0, m);  // i = 0 edge
            b.add(Y(s), Y(s), 0.5); halo(Y(s), p, 0, m, 1, 0.5);
            b.add(X(s), X(s), 0.5); halo(X(s), p, 0, m, 0, 0.5); halo(X(s), p, 0, m, 1, 0.5 * ql[m]);
            s = b.gslot(p, e, m);      // i = n-1 edge
            b.add(Y(s), Y(s), 0.5); halo(Y(s), p, 1, m, 1, 0.5);
            b.add(X(s), X(s), 0.5); halo(X(s), p, 1, m, 0, 0.5); halo(X(s), p, 1, m, 1, 0.5 * qr[m]);
            s = b.gslot(p, m, 0);      // j = 0 edge
            b.add(X(s), X(s), 0.5); halo(X(s), p, 2, m, 0, 0.5);
            b.add(Y(s), Y(s), 0.5); halo(Y(s), p, 2, m, 1, 0.5); halo(Y(s), p, 2, m, 0, 0.5 * qb[m]);
            s = b.gslot(p, m, e);      // j = n-1 edge
            b.add(X(s), X(s), 0.5); halo(X(s), p, 3, m, 0, 0.5);
            b.add(Y(s), Y(s), 0.5); halo(Y(s), p, 3, m, 1, 0.5); halo(Y(s), p, 3, m, 0, 0.5 * qt[m]);
        }
    }
    return b.nnz;
}

// ---------------------------------------------------------------------------------------------
// ECS cubic halo on xy points: interpolation weights of init_halo_xy_interp
// (equiang_cs/ecs_halo_factory_mod.f90:228-269) for target indices i = 1-hw .. n+hw and halo
// rows j = 1..hw: w[((j-1)*(n+2hw) + i-1+hw)*4 + s+1], s = -1..2, base index ind (1-based).
// init_halo_interp (ecs_halo_factory_mod.f90:193-226): target O point (t, r) of the halo row r
// beyond a panel edge at (alpha, beta) = (-pi/4 + (t-1/2) hx, pi/4 + (r-1/2) hx), projected on
// the neighbour panel, 4-point Lagrange weights about the source cell to its left
extern "C" void swmh_ecs_o_weights(int N, int rows, double *w, int *ind) {
    const double pi = pi_(), dxa = grid_hx(N);
    const int T = N + 4;
    for (int j = 1; j <= rows; ++j) {
        const double zb = 0.25 * pi + ((double)j - 0.5) * dxa;
        for (int i = -1; i <= N + 2; ++i) {
            double za = -0.25 * pi + ((double)i - 0.5) * dxa;
            double zx = std::tan(za);
            const double zz = std::tan(zb);
            zx = zx / zz;
            za = std::atan(zx);
            const int id = (int)std::floor((za + 0.25 * pi - 0.5 * dxa) / dxa) + 1;
            const double zxi = (za + 0.25 * pi) / dxa - ((double)id - 0.5);
            const int o = (j - 1) * T + (i + 1);
            ind[o] = id;
            w[4 * o + 0] = -zxi * (zxi - 1.0) * (zxi - 2.0) / 6.0;
            w[4 * o + 1] = (zxi + 1.0) * (zxi - 1.0) * (zxi - 2.0) / 2.0;
            w[4 * o + 2] = -(zxi + 1.0) * zxi * (zxi - 2.0) / 2.0;
            w[4 * o + 3] = (zxi + 1.0) * zxi * (zxi - 1.0) / 6.0;
        }
    }
}

// init_transform_matrix (ecs_halo_vec_a_factory_mod.f90:114-175): a1, a2 of the source point and
// b1, b2 of the target point in the frame of the prototype face, rotated per edge
extern "C" void swmh_ecs_vec_a_tm(int N, int rows, double *tm) {
    // rot[edge][c][k]: component c of the rotated vector is sum_k rot[edge][c][k] * v[k]
    static const int rot[4][3][3] = {{{1, 0, 0}, {0, 0, -1}, {0, 1, 0}},
                                     {{1, 0, 0}, {0, 0, 1}, {0, -1, 0}},
                                     {{0, 0, -1}, {0, 1, 0}, {1, 0, 0}},
                                     {{0, 0, 1}, {0, 1, 0}, {-1, 0, 0}}};
    static const int ab_shift[4] = {0, 0, 1, 1}, b_sign[4] = {1, -1, 1, -1};
    const double pi = pi_(), hx = grid_hx(N);
    for (int e = 0; e < 4; ++e)
        for (int j = 1; j <= rows; ++j)
            for (int i = 1; i <= N; ++i) {
                const double alpha = -0.25 * pi + ((double)i - 0.5) * hx;
                const double beta = (double)b_sign[e] * (0.25 * pi - ((double)j - 0.5) * hx);
                const double as = (double)(1 - ab_shift[e]) * alpha + (double)ab_shift[e] * beta;
                const double bs = (double)ab_shift[e] * alpha + (double)(1 - ab_shift[e]) * beta;
                auto rotate = [&](const double v[3], double out[3]) {
                    for (int c = 0; c < 3; ++c) {
                        double acc = (double)rot[e][c][0] * v[0];
                        acc = acc + (double)rot[e][c][1] * v[1];
                        acc = acc + (double)rot[e][c][2] * v[2];
                        out[c] = acc;
                    }
                };
                double v[3], a1s[3], a2s[3], r[3], b1t[3], b2t[3];
                a1_proto(as, bs, v);
                rotate(v, a1s);
                a2_proto(as, bs, v);
                rotate(v, a2s);
                r_proto(as, bs, v);
                rotate(v, r);
                const double at = std::atan(r[0] / r[2]), bt = std::atan(r[1] / r[2]);
                b1_proto(at, bt, b1t);
                b2_proto(at, bt, b2t);
                double *o = tm + (((size_t)e * rows + (j - 1)) * N + (i - 1)) * 4;
                o[0] = dot3(a1s, b1t);
                o[1] = dot3(a2s, b1t);
                o[2] = dot3(a1s, b2t);
                o[3] = dot3(a2s, b2t);
            }
}

extern "C" int swmh_ecs_xy_weights(int N, int hw, double *w, int *ind) {
    const int n = N + 1, L = n + 2 * hw;
    if (hw < 1 || n < 4) return -1;
    const double pi = pi_();
    const double dxa = grid_hx(N);
    for (int j = 1; j <= hw; ++j) {
        const double zb = 0.25 * pi + j * dxa;
        for (int i = 1 - hw; i <= n + hw; ++i) {
            double za = -0.25 * pi + (i - 1.0) * dxa;
            // target point on the prototype face, projected through the sphere onto the
            // source face: alpha' = atan(tan(alpha) / tan(beta))
            double zx = std::tan(za);
            const double zz = std::tan(zb);
            zx = zx / zz;
            za = std::atan(zx);
            int id = (int)std::floor((za + 0.25 * pi) / dxa) + 1;
            id = std::min(std::max(2, id), n - 2);
            const double t = (za + 0.25 * pi) / dxa - (id - 1.0);
            const int o = (j - 1) * L + (i - 1 + hw);
            ind[o] = id;
            // cubic Lagrange weights on the nodes -1, 0, 1, 2
            w[4 * o + 0] = -t * (t - 1.0) * (t - 2.0) / 6.0;
            w[4 * o + 1] = (t + 1.0) * (t - 1.0) * (t - 2.0) / 2.0;
            w[4 * o + 2] = -(t + 1.0) * t * (t - 2.0) / 2.0;
            w[4 * o + 3] = (t + 1.0) * t * (t - 1.0) / 6.0;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// C grid (staggering 'C'): metric terms at the points of one type and the cube-edge maps of its
// halos, both from the Ah machinery above on the same sphere.
//
// Points of type t (mesh_factory_mod.f90:34-47): alpha = alpha0 + (i-1+si)*hx, beta likewise with
// sj; o (t=0): si = sj = 1/2, x (t=1): si = 0, sj = 1/2, y (t=2): si = 1/2, sj = 0, q (t=3): 0, 0.
extern "C" void swmh_c_geometry(int N, int t, int H, int rows, int pitch, double omega, const double axis[3],
                                double *J, double *Qi11, double *Qi12, double *Qi22, double *f) {
    const double si = (t == 0 || t == 2) ? 0.5 : 0.0, sj = (t == 0 || t == 1) ? 0.5 : 0.0;
    const double hx = grid_hx(N), alpha0 = -0.25 * pi_();
    for (int p = 0; p < 6; ++p)
        for (int j = 1; j <= N + 1; ++j) {
            const double beta = alpha0 + ((double)(j - 1) + sj) * hx;
            for (int i = 1; i <= N + 1; ++i) {
                const double alpha = alpha0 + ((double)(i - 1) + si) * hx;
                const long id = ((long)p * rows + (j - 1 + H)) * pitch + (i - 1 + H);
                const double ta = std::tan(alpha), tb = std::tan(beta);
                const double sg = std::sqrt(1.0 + ta * ta + tb * tb);
                if (J) J[id] = (1.0 + ta * ta) * (1.0 + tb * tb) / (sg * sg * sg);
                if (Qi11) Qi11[id] = sg * sg / (1.0 + ta * ta);
                if (Qi12) Qi12[id] = sg * sg * ta * tb / ((1 + ta * ta) * (1 + tb * tb));
                if (Qi22) Qi22[id] = sg * sg / (1.0 + tb * tb);
                if (f) {
                    double rp[3], rr[3];
                    r_proto(alpha, beta, rp);
                    to_panel(p, rp, rr);
                    const double pr = (axis[0] * rr[0] + axis[1] * rr[1] + axis[2] * rr[2]) /
                                      std::sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2]);
                    f[id] = 2 * omega * pr;
                }
            }
        }
}

// The C points of one panel side lie on the Ah lattice of 2N cells (doubled coordinates
// 0..2N: faces at even, centres at odd values).  A point of a C halo at doubled depth d2 beyond
// edge e of panel p, at doubled position T2 along the edge, is entry (k = T2, d = d2) of the Ah
// neighbour table of that doubled lattice: the neighbour panel q across e, the doubled
// coordinates (a, b) there, and the vector code of the components.  Face points normal to the
// edge start at d2 = 0 (halo index 0 holds the neighbour's copy of the shared face:
// exchange_factory_mod.f90:402-403,420-423 widen the exchange of those points by one), cell
// centres at d2 = 1 (halo index 0 is the neighbour's centre nearest the edge).
//
// Entries of the gather lists are element offsets of the point in the stored C field layout
// (+ fs for the second component of a vector pair); signs are +-1.
namespace {
long c_off(int H, int rows, int pitch, int p, int i, int j) {
    return ((long)p * rows + (j - 1 + H)) * pitch + (i - 1 + H);
}
}  // namespace

extern "C" int swmh_c_halo_lists(int N, int H, int rows, int pitch, long long fs, int depth_vec, int depth_o,
                                 long long *vec_dst, long long *vec_src, signed char *vec_sign,
                                 long long *face_dst, long long *face_src, signed char *face_sign,
                                 long long *o_dst, long long *o_src, long long *q_own, long long *q_nb,
                                 int *counts) {
    const int N2 = 2 * N, n2 = N2 + 1, nd = 2 * std::max(depth_vec, depth_o);
    if (depth_vec < 1 || depth_o < 1 || depth_vec >= H || depth_o >= H) return -1;
    std::vector<int> tab((size_t)6 * 4 * n2 * nd * 4);
    swmh_halo_table(N2, nd, tab.data());
    auto entry = [&](int p, int e, int k2, int d2) { return &tab[((((size_t)p * 4 + e) * n2 + k2) * nd + d2) * 4]; };
    int nv = 0, nf = 0, no = 0, nq = 0;
    for (int p = 0; p < 6; ++p)
        for (int e = 0; e < 4; ++e) {
            const bool xn = e < 2;  // the edge's normal is the x axis
            // vector components: comp 0 = u (x points), 1 = v (y points); tangential ones
            // (u beyond y edges, v beyond x edges) to depth_vec, the face ones to depth 1
            for (int comp = 0; comp < 2; ++comp) {
                const bool faces = (xn && comp == 0) || (!xn && comp == 1);
                const int ntan = faces ? N : N + 1;
                const int depth = faces ? 1 : depth_vec;
                for (int d = 1; d <= depth; ++d)
                    for (int k = 1; k <= ntan; ++k) {
                        const int d2 = faces ? 2 * (d - 1) : 2 * d - 1;
                        const int T2 = faces ? 2 * k - 1 : 2 * (k - 1);
                        const int *t = entry(p, e, T2, d2);
                        const int q = t[0], a = t[1], b = t[2], code = t[3];
                        const bool swap = code & 1;
                        const int qcomp = swap ? 1 - comp : comp;
                        const bool neg = comp == 0 ? (code & 2) : (code & 4);
                        // the neighbour's point of component qcomp: faces even along its axis
                        const bool ok = qcomp == 0 ? (a % 2 == 0 && b % 2 == 1) : (a % 2 == 1 && b % 2 == 0);
                        if (!ok) return -2;
                        const int qi = qcomp == 0 ? a / 2 + 1 : (a + 1) / 2, qj = qcomp == 0 ? (b + 1) / 2 : b / 2 + 1;
                        int i, j;
                        const int nx = N + (comp == 0), ny = N + (comp == 1);
                        if (e == 0) { i = 1 - d; j = k; }
                        else if (e == 1) { i = nx + d; j = k; }
                        else if (e == 2) { i = k; j = 1 - d; }
                        else { i = k; j = ny + d; }
                        const long long dst = comp * fs + c_off(H, rows, pitch, p, i, j);
                        const long long src = qcomp * fs + c_off(H, rows, pitch, q, qi, qj);
                        if (faces) {
                            face_dst[nf] = dst; face_src[nf] = src; face_sign[nf] = neg ? -1 : 1; ++nf;
                        } else {
                            vec_dst[nv] = dst; vec_src[nv] = src; vec_sign[nv] = neg ? -1 : 1; ++nv;
                        }
                    }
            }
            // O points (cell centres), cross halo to depth_o
            for (int d = 1; d <= depth_o; ++d)
                for (int k = 1; k <= N; ++k) {
                    const int *t = entry(p, e, 2 * k - 1, 2 * d - 1);
                    const int q = t[0], a = t[1], b = t[2];
                    if (a % 2 != 1 || b % 2 != 1) return -3;
                    int i, j;
                    if (e == 0) { i = 1 - d; j = k; }
                    else if (e == 1) { i = N + d; j = k; }
                    else if (e == 2) { i = k; j = 1 - d; }
                    else { i = k; j = N + d; }
                    o_dst[no] = c_off(H, rows, pitch, p, i, j);
                    o_src[no] = c_off(H, rows, pitch, q, (a + 1) / 2, (b + 1) / 2);
                    ++no;
                }
            // q points on the edge: the neighbour's copy of each (k = 1..N+1)
            for (int k = 1; k <= N + 1; ++k) {
                const int *t = entry(p, e, 2 * (k - 1), 0);
                const int q = t[0], a = t[1], b = t[2];
                if (a % 2 != 0 || b % 2 != 0) return -4;
                int i, j;
                if (e == 0) { i = 1; j = k; }
                else if (e == 1) { i = N + 1; j = k; }
                else if (e == 2) { i = k; j = 1; }
                else { i = k; j = N + 1; }
                q_own[nq] = c_off(H, rows, pitch, p, i, j);
                q_nb[nq] = c_off(H, rows, pitch, q, a / 2 + 1, b / 2 + 1);
                ++nq;
            }
        }
    counts[0] = nv; counts[1] = nf; counts[2] = no; counts[3] = nq;
    return 0;
}

namespace {
// init_1d_tile_interp (cubic_lag, ecs_halo_vec_c_factory_mod.f90:272-341) for positions
// 1..npos and halo rows 1..hw: source grid shift sx, target shift (tx, ty), stencil base clamped
// to [2, nx-2]
void c_ecs_interp(int N, int npos, int hw, int nx, double sx, double tx, double ty, double *w, int *ind) {
    const double pi = pi_(), hx = grid_hx(N);
    for (int j = 1; j <= hw; ++j) {
        const double beta = 0.25 * pi + ((double)j + ty) * hx;
        for (int i = 1; i <= npos; ++i) {
            double alpha = -0.25 * pi + ((double)i - 1.0 + tx) * hx;
            double x = std::tan(alpha), y = std::tan(beta);
            const double z = 1.0 / std::sqrt(1.0 + x * x + y * y);
            x = x * z;
            y = y * z;
            alpha = std::atan(x / y);
            double xi = (alpha + 0.25 * pi - sx * hx) / hx;
            const int id = std::max(2, std::min((int)xi + 1, nx - 2));
            xi = xi - id + 1;
            const int o = (i - 1) * hw + (j - 1);
            ind[o] = id;
            w[4 * o + 0] = -xi * (xi - 1.0) * (xi - 2.0) / 6.0;
            w[4 * o + 1] = (xi + 1.0) * (xi - 1.0) * (xi - 2.0) / 2.0;
            w[4 * o + 2] = -(xi + 1.0) * xi * (xi - 2.0) / 2.0;
            w[4 * o + 3] = (xi + 1.0) * xi * (xi - 1.0) / 6.0;
        }
    }
}
// the halo point's coordinate along the edge in the neighbour's frame
double c_ecs_along(double alpha, double beta) {
    double x = std::tan(alpha), y = std::tan(beta);
    const double z = 1.0 / std::sqrt(1.0 + x * x + y * y);
    x = x * z;
    y = y * z;
    return std::atan(x / y);
}
}  // namespace

extern "C" int swmh_c_ecs_vec_tables(int N, int hw, int covariant, double *w0, int *i0, double *w1, int *i1,
                                     double *w2, int *i2, double *q1, double *q2, double *ct) {
    if (hw < 1 || hw > 5 || N < 8) return -1;
    const double pi = pi_(), hx = grid_hx(N);
    c_ecs_interp(N, N + 1, hw, N + 1, 0.0, 0.0, -0.5, w0, i0);   // tangential, "x" -> "x"
    c_ecs_interp(N, N, hw, N, 0.5, 0.5, 0.0, w1, i1);            // normal, "y" -> "y"
    if (covariant) c_ecs_interp(N, N + 1, hw, N, 0.0, 0.5, 0.0, w2, i2);   // "x" -> "y"
    else c_ecs_interp(N, N + 1, hw, N, 0.5, 0.0, -0.5, w2, i2);            // "y" -> "x"
    for (int j = 1; j <= hw; ++j)
        for (int i = 1; i <= N + 1; ++i) {
            const int o = (i - 1) * hw + (j - 1);
            double v[3], a[3], b[3], e[3];
            if (!covariant) {
                // init_basis_transform (:343-380): qtt, qtn
                const double beta = 0.25 * pi + ((double)j - 0.5) * hx, alpha = -0.25 * pi + ((double)i - 1.0) * hx;
                const double al1 = c_ecs_along(alpha, beta), be1 = -0.25 * pi + ((double)j - 0.5) * hx;
                b1_proto(alpha, beta, b);
                a1_proto(al1, be1, v);
                a[0] = v[0], a[1] = v[2], a[2] = -v[1];
                a2_proto(al1, be1, v);
                e[0] = v[0], e[1] = v[2], e[2] = -v[1];
                q1[o] = dot3(a, b);
                q2[o] = dot3(e, b);
            } else {
                // init_basis_transform_cov (:382-419): qnn first, then qnt
                const double beta = 0.25 * pi + (double)j * hx, alpha = -0.25 * pi + ((double)i - 0.5) * hx;
                const double al1 = c_ecs_along(alpha, beta), be1 = -0.25 * pi + (double)j * hx;
                a2_proto(alpha, beta, a);
                b1_proto(al1, be1, v);
                b[0] = v[0], b[1] = v[2], b[2] = -v[1];
                b2_proto(al1, be1, v);
                e[0] = v[0], e[1] = v[2], e[2] = -v[1];
                q1[o] = dot3(a, e);
                q2[o] = dot3(a, b);
                // init_tangent_comp_stretch (:421-449)
                const double bt = 0.25 * pi + ((double)j - 0.5) * hx, at = -0.25 * pi + ((double)i - 1.0) * hx;
                const double at1 = c_ecs_along(at, bt), bt1 = -0.25 * pi + ((double)j - 0.5) * hx;
                double tv[3], sv[3];
                a1_proto(at, bt, tv);
                a1_proto(at1, bt1, sv);
                ct[o] = std::sqrt((tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]) /
                                  (sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]));
            }
        }
    return 0;
}

extern "C" int swmh_c_ecs_strip_sources(int N, int H, int rows, int pitch, long long fs, int depth,
                                        long long *nsrc, signed char *nsign, long long *tsrc, signed char *tsign,
                                        long long *osrc) {
    const int N2 = 2 * N, n2 = N2 + 1, nd = 2 * depth;
    if (depth < 1 || nd > n2) return -1;
    std::vector<int> tab((size_t)6 * 4 * n2 * nd * 4);
    swmh_halo_table(N2, nd, tab.data());
    auto entry = [&](int p, int e, int k2, int d2) { return &tab[((((size_t)p * 4 + e) * n2 + k2) * nd + d2) * 4]; };
    for (int p = 0; p < 6; ++p)
        for (int e = 0; e < 4; ++e) {
            const int normal = e < 2 ? 0 : 1;   // the component normal to the edge: u at x edges
            for (int D = 0; D < depth; ++D) {
                // normal component: faces at depth D (0 = the neighbour's copy of the shared face)
                for (int k = 1; k <= N; ++k) {
                    const int *t = entry(p, e, 2 * k - 1, 2 * D);
                    const int q = t[0], a = t[1], b = t[2], code = t[3];
                    const int qc = (code & 1) ? 1 - normal : normal;
                    const bool neg = normal == 0 ? (code & 2) : (code & 4);
                    if (qc == 0 ? (a % 2 != 0 || b % 2 != 1) : (a % 2 != 1 || b % 2 != 0)) return -2;
                    const int qi = qc == 0 ? a / 2 + 1 : (a + 1) / 2, qj = qc == 0 ? (b + 1) / 2 : b / 2 + 1;
                    const size_t o = (((size_t)p * 4 + e) * depth + D) * N + (k - 1);
                    nsrc[o] = qc * fs + c_off(H, rows, pitch, q, qi, qj);
                    nsign[o] = neg ? -1 : 1;
                }
                // tangential component: centre lines at depth D
                const int tc = 1 - normal;
                for (int k = 1; k <= N + 1; ++k) {
                    const int *t = entry(p, e, 2 * (k - 1), 2 * D + 1);
                    const int q = t[0], a = t[1], b = t[2], code = t[3];
                    const int qc = (code & 1) ? 1 - tc : tc;
                    const bool neg = tc == 0 ? (code & 2) : (code & 4);
                    if (qc == 0 ? (a % 2 != 0 || b % 2 != 1) : (a % 2 != 1 || b % 2 != 0)) return -3;
                    const int qi = qc == 0 ? a / 2 + 1 : (a + 1) / 2, qj = qc == 0 ? (b + 1) / 2 : b / 2 + 1;
                    const size_t o = (((size_t)p * 4 + e) * depth + D) * (N + 1) + (k - 1);
                    tsrc[o] = qc * fs + c_off(H, rows, pitch, q, qi, qj);
                    tsign[o] = neg ? -1 : 1;
                }
                // cell centres at depth D
                for (int k = 1; k <= N; ++k) {
                    const int *t = entry(p, e, 2 * k - 1, 2 * D + 1);
                    const int q = t[0], a = t[1], b = t[2];
                    if (a % 2 != 1 || b % 2 != 1) return -4;
                    osrc[(((size_t)p * 4 + e) * depth + D) * N + (k - 1)] = c_off(H, rows, pitch, q, (a + 1) / 2, (b + 1) / 2);
                }
            }
        }
    return 0;
}
