// Cross-check for the oracle's 6x6 Cholesky: the reference solves its normal
// equations with Eigen::LLT<Eigen::Matrix<float,6,6>> (project/src/
// linear_least_squares.cpp:41, Eigen 3.3.7 vendored at 3rd/eigen).  This file
// compiles that Eigen routine, from the reference's own vendored headers, into
// oracle/_ref/libref_llt.so.  Test infrastructure only.
#include <Eigen/Dense>

extern "C" int ref_llt_solve6_f32(const float* ata_colmajor, const float* atb, float* x)
{
    Eigen::Matrix<float, 6, 6> A;
    Eigen::Matrix<float, 6, 1> b;
    for (int c = 0; c < 6; ++c)
        for (int r = 0; r < 6; ++r) A(r, c) = ata_colmajor[c * 6 + r];
    for (int r = 0; r < 6; ++r) b(r) = atb[r];
    Eigen::LLT<Eigen::Matrix<float, 6, 6>> llt(A);
    Eigen::Matrix<float, 6, 1> s = llt.solve(b);
    for (int r = 0; r < 6; ++r) x[r] = s(r);
    return llt.info() == Eigen::Success ? 0 : 1;
}

extern "C" int ref_llt_solve6_f64(const double* ata_colmajor, const double* atb, double* x)
{
    Eigen::Matrix<double, 6, 6> A;
    Eigen::Matrix<double, 6, 1> b;
    for (int c = 0; c < 6; ++c)
        for (int r = 0; r < 6; ++r) A(r, c) = ata_colmajor[c * 6 + r];
    for (int r = 0; r < 6; ++r) b(r) = atb[r];
    Eigen::LLT<Eigen::Matrix<double, 6, 6>> llt(A);
    Eigen::Matrix<double, 6, 1> s = llt.solve(b);
    for (int r = 0; r < 6; ++r) x[r] = s(r);
    return llt.info() == Eigen::Success ? 0 : 1;
}
