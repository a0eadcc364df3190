// SBP first-derivative operators on Ah (corner) points and their quadrature weights.
//
// The coefficient tables are the reference's (differential_operators/sbp_operators/
// sbp_operators_collection_mod.f90: Q21/Da2_in :15-20, Q42/Da4_in :23-30,84-85,
// Q43 :73-81, Q63/Da6_in :33-67, Q21_A/Q42_A/Q63_A :17,31,33-35); they define the
// discretisation and are reproduced value for value.  The right-boundary block is the
// left block reflected in both indices with the sign flipped, as built by
// create_sbp_operator (sbp_factory_mod.f90:162-165).
//
// All tables are constexpr so that host code (quadrature) and device code (stencils)
// see identical double values; FP32 kernels round them once at compile time.
#pragma once

#if defined(__HIPCC__)
#define SWM_HD __host__ __device__
#else
#define SWM_HD
#endif

namespace swm {

template <int ORD> struct Sbp;

// d21: left row f2-f1, interior (f(i+1)-f(i-1))/2
template <> struct Sbp<21> {
    static constexpr int NW = 2, NE = 1, WIN = 3, INSHIFT = -1, REACH = 1;
    static constexpr int last[1] = {2};
    static constexpr double Ql[1][2] = {{-1.0, 1.0}};
    static constexpr double Qin[3] = {-0.5, 0.0, 0.5};
};

// d42: 4-row boundary closure, 4th-order interior
template <> struct Sbp<42> {
    static constexpr int NW = 6, NE = 4, WIN = 5, INSHIFT = -2, REACH = 3;
    static constexpr int last[4] = {4, 3, 5, 6};
    static constexpr double Ql[4][6] = {
        {-1.4117647058823528, 1.7352941176470589, -0.2352941176470588, -0.08823529411764705, 0.0, 0.0},
        {-0.5, 0.0, 0.5, 0.0, 0.0, 0.0},
        {0.09302325581395347, -0.686046511627907, 0.0, 0.686046511627907, -0.09302325581395347, 0.0},
        {0.030612244897959186, -0.0, -0.6020408163265307, 0.0, 0.653061224489796, -0.0816326530612245}};
    static constexpr double Qin[5] = {1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0};
};

// d43: 4-row closure with full 6-point rows, 4th-order interior
template <> struct Sbp<43> {
    static constexpr int NW = 6, NE = 4, WIN = 5, INSHIFT = -2, REACH = 5;
    static constexpr int last[4] = {6, 6, 6, 6};
    static constexpr double Ql[4][6] = {
        {-1.8280236842718398, 2.978054584697369, -1.4653148294044123, 0.3078538227474199, 0.008136925288119662, -0.0007068190566565205},
        {-0.37814159126691327, -0.31480210500706657, 0.707290999364066, 0.04835554461933453, -0.06866771096803422, 0.005964863258613564},
        {0.1162478575921292, -0.8028500807354999, 0.21558841368737428, 0.5078566674295846, -0.03231754093993847, -0.004525317033649745},
        {-0.010287398573128293, 0.12592344801071614, -0.7341531396449149, 0.04979271660173103, 0.6506171865540596, -0.08189281294846368}};
    static constexpr double Qin[5] = {1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0};
};

// d63: 6-row closure, 6th-order interior; entries are the reference's rational expressions
namespace d63c {
constexpr double A1 = 0.13649e5 / 0.43200e5, A2 = 0.12013e5 / 0.8640e4, A3 = 0.2711e4 / 0.4320e4,
                 A4 = 0.5359e4 / 0.4320e4, A5 = 0.7877e4 / 0.8640e4, A6 = 0.43801e5 / 0.43200e5;
constexpr double q56 = 5591070156686698065364559.0 / 7931626489314500743872000.0;
constexpr double q11 = -1.0 / 2.0;
constexpr double q12 = -0.953e3 / 0.16200e5 + q56;
constexpr double q13 = 0.715489e6 / 0.259200e6 - (4.0 * q56);
constexpr double q14 = -0.62639e5 / 0.14400e5 + (6.0 * q56);
constexpr double q15 = 0.147127e6 / 0.51840e5 - (4.0 * q56);
constexpr double q16 = -0.89387e5 / 0.129600e6 + q56;
constexpr double q23 = -0.57139e5 / 0.8640e4 + (10.0 * q56);
constexpr double q24 = 0.745733e6 / 0.51840e5 - (20.0 * q56);
constexpr double q25 = -0.18343e5 / 0.1728e4 + (15.0 * q56);
constexpr double q26 = 0.240569e6 / 0.86400e5 - (4.0 * q56);
constexpr double q34 = -0.176839e6 / 0.12960e5 + (20.0 * q56);
constexpr double q35 = 0.242111e6 / 0.17280e5 - (20.0 * q56);
constexpr double q36 = -0.182261e6 / 0.43200e5 + (6.0 * q56);
constexpr double q45 = -0.165041e6 / 0.25920e5 + (10.0 * q56);
constexpr double q46 = 0.710473e6 / 0.259200e6 - (4.0 * q56);
constexpr double q47 = 1.0 / 6.0e1, q57 = -3.0 / 2.0e1, q58 = 1.0 / 6.0e1;
constexpr double q67 = 3.0 / 4.0, q68 = -3.0 / 2.0e1, q69 = 1.0 / 6.0e1;
}  // namespace d63c

template <> struct Sbp<63> {
    static constexpr int NW = 9, NE = 6, WIN = 7, INSHIFT = -3, REACH = 5;
    static constexpr int last[6] = {6, 6, 6, 7, 8, 9};
    static constexpr double Ql[6][9] = {
        {d63c::q11 / d63c::A1, d63c::q12 / d63c::A1, d63c::q13 / d63c::A1, d63c::q14 / d63c::A1, d63c::q15 / d63c::A1, d63c::q16 / d63c::A1, 0.0, 0.0, 0.0},
        {-d63c::q12 / d63c::A2, 0.0, d63c::q23 / d63c::A2, d63c::q24 / d63c::A2, d63c::q25 / d63c::A2, d63c::q26 / d63c::A2, 0.0, 0.0, 0.0},
        {-d63c::q13 / d63c::A3, -d63c::q23 / d63c::A3, 0.0, d63c::q34 / d63c::A3, d63c::q35 / d63c::A3, d63c::q36 / d63c::A3, 0.0, 0.0, 0.0},
        {-d63c::q14 / d63c::A4, -d63c::q24 / d63c::A4, -d63c::q34 / d63c::A4, 0.0, d63c::q45 / d63c::A4, d63c::q46 / d63c::A4, d63c::q47 / d63c::A4, 0.0, 0.0},
        {-d63c::q15 / d63c::A5, -d63c::q25 / d63c::A5, -d63c::q35 / d63c::A5, -d63c::q45 / d63c::A5, 0.0, d63c::q56 / d63c::A5, d63c::q57 / d63c::A5, d63c::q58 / d63c::A5, 0.0},
        {-d63c::q16 / d63c::A6, -d63c::q26 / d63c::A6, -d63c::q36 / d63c::A6, -d63c::q46 / d63c::A6, -d63c::q56 / d63c::A6, 0.0, d63c::q67 / d63c::A6, d63c::q68 / d63c::A6, d63c::q69 / d63c::A6}};
    static constexpr double Qin[7] = {-1.0 / 6.0e1, 3.0 / 2.0e1, -3.0 / 4.0, 0.0, 3.0 / 4.0, -3.0 / 2.0e1, 1.0 / 6.0e1};
};

// SBP quadrature edge weights (Ah points), mirrored at the far end
struct QuadEdge {
    int ne;
    double a[6];
};
inline QuadEdge quad_edge(int order) {
    if (order == 21) return {2, {0.5, 1.0}};
    if (order == 42) return {4, {17.0 / 48.0, 59.0 / 48.0, 43.0 / 48.0, 49.0 / 48.0}};
    if (order == 63)
        return {6, {0.13649e5 / 0.43200e5, 0.12013e5 / 0.8640e4, 0.2711e4 / 0.4320e4, 0.5359e4 / 0.4320e4,
                    0.7877e4 / 0.8640e4, 0.43801e5 / 0.43200e5}};
    return {0, {}};
}

}  // namespace swm
