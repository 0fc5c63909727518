// gf_eigen.hpp -- per-lane FP64 symmetric 3x3 eigensolver and the stage
// output formulas, as device functions.
//
// The solver follows Eigen 3.4's SelfAdjointEigenSolver<Matrix3d>::compute
// (the solver PDAL's three stages call): scale by the largest lower-triangle
// magnitude, the closed 3x3 Householder tridiagonalization, implicit
// symmetric QR steps with Wilkinson shift (at most 30*n sweeps), ascending
// selection sort of values and vector columns
// (vendor/eigen/Eigen/src/Eigenvalues/SelfAdjointEigenSolver.h:420-575,
// 839-890; Tridiagonalization.h:464-504; Jacobi/Jacobi.h:231-266).  With
// contraction off it reproduces the host solver's bits, so eigenvector signs
// (which decide normals whose z component is exactly 0) agree with PDAL.
#pragma once

#include "gf_internal.hpp"

namespace gf {

__device__ __forceinline__ void eig_givens(double p, double q, double& c, double& s)
{
    if (q == 0.0) {
        c = p < 0.0 ? -1.0 : 1.0;
        s = 0.0;
    } else if (p == 0.0) {
        c = 0.0;
        s = q < 0.0 ? 1.0 : -1.0;
    } else if (fabs(p) > fabs(q)) {
        double t = q / p;
        double u = sqrt(1.0 + t * t);
        if (p < 0.0) u = -u;
        c = 1.0 / u;
        s = -t * c;
    } else {
        double t = p / q;
        double u = sqrt(1.0 + t * t);
        if (q < 0.0) u = -u;
        s = -1.0 / u;
        c = -t * s;
    }
}

__device__ __forceinline__ double eig_hypot(double x, double y)
{
    x = fabs(x);
    y = fabs(y);
    if (isinf(x) || isinf(y)) return INFINITY;
    if (isnan(x) || isnan(y)) return NAN;
    double p = x < y ? y : x;
    double m = y < x ? y : x;
    if (p == 0.0) return 0.0;
    double qp = m / p;
    return p * sqrt(1.0 + qp * qp);
}

// cov: xx, xy, xz, yy, yz, zz.  Returns 0 on success, 1 on NoConvergence.
// ev ascending; vec column-major (vec[3*j + axis] belongs to ev[j]).
__device__ inline int eigen_sym3(const double cov[6], double ev[3], double vec[9])
{
    double m00 = cov[0], m10 = cov[1], m20 = cov[2], m11 = cov[3], m21 = cov[4],
           m22 = cov[5];
    if (isnan(m00) || isnan(m10) || isnan(m20) || isnan(m11) || isnan(m21) ||
        isnan(m22)) {
        for (int i = 0; i < 3; ++i) ev[i] = NAN;
        for (int i = 0; i < 9; ++i) vec[i] = NAN;
        return 1;
    }
    double scale = 0.0;
    {
        double lo[6] = {m00, m10, m11, m20, m21, m22};
        for (int i = 0; i < 6; ++i)
            if (fabs(lo[i]) > scale) scale = fabs(lo[i]);
    }
    if (scale == 0.0) scale = 1.0;
    m00 /= scale;
    m10 /= scale;
    m11 /= scale;
    m20 /= scale;
    m21 /= scale;
    m22 /= scale;

    double d[3], e[2], q[9];
    d[0] = m00;
    double v1n2 = m20 * m20;
    if (v1n2 <= 2.2250738585072014e-308) {
        d[1] = m11;
        d[2] = m22;
        e[0] = m10;
        e[1] = m21;
        q[0] = 1; q[1] = 0; q[2] = 0;
        q[3] = 0; q[4] = 1; q[5] = 0;
        q[6] = 0; q[7] = 0; q[8] = 1;
    } else {
        double beta = sqrt(m10 * m10 + v1n2);
        double inv = 1.0 / beta;
        double m01 = m10 * inv;
        double m02 = m20 * inv;
        double qq = 2.0 * m01 * m21 + m02 * (m22 - m11);
        d[1] = m11 + m02 * qq;
        d[2] = m22 - m02 * qq;
        e[0] = beta;
        e[1] = m21 - m01 * qq;
        q[0] = 1; q[3] = 0;   q[6] = 0;
        q[1] = 0; q[4] = m01; q[7] = m02;
        q[2] = 0; q[5] = m02; q[8] = -m01;
    }

    const double tiny = 2.2250738585072014e-308;
    const double prec_inv = 1.0 / 2.220446049250313e-16;
    int end = 2, start = 0, iter = 0;
    while (end > 0) {
        for (int i = start; i < end; ++i) {
            if (fabs(e[i]) < tiny) {
                e[i] = 0.0;
            } else {
                double sc = prec_inv * e[i];
                if (sc * sc <= fabs(d[i]) + fabs(d[i + 1])) e[i] = 0.0;
            }
        }
        while (end > 0 && e[end - 1] == 0.0) end--;
        if (end <= 0) break;
        iter++;
        if (iter > 90) break;
        start = end - 1;
        while (start > 0 && e[start - 1] != 0.0) start--;

        // Francis step with Wilkinson shift on rows start..end.
        double td = (d[end - 1] - d[end]) * 0.5;
        double ee = e[end - 1];
        double mu = d[end];
        if (td == 0.0) {
            mu -= fabs(ee);
        } else if (ee != 0.0) {
            double e2 = ee * ee;
            double h = eig_hypot(td, ee);
            if (e2 == 0.0)
                mu -= ee / ((td + (td > 0.0 ? h : -h)) / ee);
            else
                mu -= e2 / (td + (td > 0.0 ? h : -h));
        }
        double xx = d[start] - mu;
        double zz = e[start];
        for (int k = start; k < end && zz != 0.0; ++k) {
            double c, s;
            eig_givens(xx, zz, c, s);
            double sdk = s * d[k] + c * e[k];
            double dkp1 = s * e[k] + c * d[k + 1];
            d[k] = c * (c * d[k] - s * e[k]) - s * (c * e[k] - s * d[k + 1]);
            d[k + 1] = s * sdk + c * dkp1;
            e[k] = c * sdk - s * dkp1;
            if (k > start) e[k - 1] = c * e[k - 1] - s * zz;
            xx = e[k];
            if (k < end - 1) {
                zz = -s * e[k + 1];
                e[k + 1] = c * e[k + 1];
            }
            for (int i = 0; i < 3; ++i) {
                double xi = q[3 * k + i], yi = q[3 * (k + 1) + i];
                q[3 * k + i] = c * xi + (-s) * yi;
                q[3 * (k + 1) + i] = s * xi + c * yi;
            }
        }
    }
    int info = iter <= 90 ? 0 : 1;
    if (info == 0) {
        for (int i = 0; i < 2; ++i) {
            int k = 0;
            double best = d[i];
            for (int j = 1; j < 3 - i; ++j)
                if (d[i + j] < best) {
                    best = d[i + j];
                    k = j;
                }
            if (k > 0) {
                double t = d[i];
                d[i] = d[i + k];
                d[i + k] = t;
                for (int r = 0; r < 3; ++r) {
                    double u = q[3 * i + r];
                    q[3 * i + r] = q[3 * (i + k) + r];
                    q[3 * (i + k) + r] = u;
                }
            }
        }
    }
    for (int i = 0; i < 3; ++i) ev[i] = d[i] * scale;
    for (int i = 0; i < 9; ++i) vec[i] = q[i];
    return info;
}

// Eigen::isZero with the double dummy precision: every |coefficient| <= 1e-12.
__device__ __forceinline__ bool cov_is_zero(const double c[6])
{
    for (int i = 0; i < 6; ++i)
        if (!(fabs(c[i]) <= 1e-12)) return false;
    return true;
}

// Writes one stage's columns for one row; returns the row status.
// Formulas: filters/CovarianceFeaturesFilter.cpp:249-365 (covfeat),
// pdal/private/MathUtils.cpp:589-608 + MathUtils.hpp:321-324 (normal),
// filters/EigenvaluesFilter.cpp:152-167 (eigenvalues).
__device__ inline uint8_t stage_outputs(const StageProgram& sp, int k,
                                        const double cov[6], bool zero,
                                        int info, const double ev_in[3],
                                        const double vec[9], double* vals)
{
    if (sp.stage == STAGE_NORMAL) {
        if (k < 3) return ST_FEW_NEIGHBORS;
        if (zero) return ST_ZERO_COV;
        if (info) return ST_EIGEN_FAIL;
        double sum = ev_in[0] + ev_in[1] + ev_in[2];
        double curv = sum != 0.0 ? fabs(ev_in[0] / sum) : 0.0;
        double nx = vec[0], ny = vec[1], nz = vec[2];
        if (sp.always_up && nz < 0.0) {
            nx = -nx;
            ny = -ny;
            nz = -nz;
        }
        vals[0] = nx;
        vals[1] = ny;
        vals[2] = nz;
        vals[3] = curv;
        return 0;
    }
    if (sp.stage == STAGE_RANK) {
        // math::computeRank (pdal/private/MathUtils.cpp:197-208): Eigen's
        // JacobiSVD::rank() with the threshold narrowed to float -- the
        // singular values of the symmetric covariance are its eigenvalues'
        // magnitudes; count those (nonzero) at or above max(thresh * largest,
        // DBL_MIN).  No isZero skip: a zero matrix has rank 0.
        if (info) return ST_EIGEN_FAIL;
        double a = fabs(ev_in[0]), b = fabs(ev_in[1]), c = fabs(ev_in[2]);
        double t;
        if (a < b) { t = a; a = b; b = t; }
        if (b < c) { t = b; b = c; c = t; }
        if (a < b) { t = a; a = b; b = t; }
        const double thr = static_cast<double>(static_cast<float>(sp.p1));
        const double pre = fmax(a * thr, 2.2250738585072014e-308);
        int r = 0;
        if (a != 0.0 && a >= pre) r = 1;
        if (r == 1 && b != 0.0 && b >= pre) r = 2;
        if (r == 2 && c != 0.0 && c >= pre) r = 3;
        vals[0] = static_cast<double>(r);
        return 0;
    }
    if (zero) return ST_ZERO_COV;
    if (info) return ST_EIGEN_FAIL;
    if (sp.stage == STAGE_COPLANAR) {
        // ApproximateCoplanarFilter.cpp:117-121, eigenvalues ascending.
        vals[0] = (ev_in[1] > sp.p1 * ev_in[0] && sp.p2 * ev_in[1] > ev_in[2]) ? 1.0 : 0.0;
        return 0;
    }
    if (sp.stage == STAGE_EIGENVALUES) {
        double e0 = ev_in[0], e1 = ev_in[1], e2 = ev_in[2];
        if (sp.normalize) {
            double sum = e0 + e1 + e2;
            e0 = e0 / sum;
            e1 = e1 / sum;
            e2 = e2 / sum;
        }
        vals[0] = e0;
        vals[1] = e1;
        vals[2] = e2;
        return 0;
    }
    double lam[3];
    for (int i = 0; i < 3; ++i) {
        double e = ev_in[2 - i];
        lam[i] = (e < 0.0) ? 0.0 : e;
    }
    double sum = 0.0;
    sum = sum + lam[0];
    sum = sum + lam[1];
    sum = sum + lam[2];
    if (lam[0] == 0.0) return ST_FEATURE_INVALID;
    if (sp.mode == MODE_SQRT) {
        for (int i = 0; i < 3; ++i) lam[i] = sqrt(lam[i]);
    } else if (sp.mode == MODE_NORMALIZED) {
        for (int i = 0; i < 3; ++i) lam[i] = lam[i] / sum;
    }
    int c = 0;
    for (int b = 0; b < 10; ++b) {
        if (!(sp.features & (1u << b))) continue;
        double v = 0.0;
        switch (b) {
        case 0: v = (lam[0] - lam[1]) / lam[0]; break;
        case 1: v = (lam[1] - lam[2]) / lam[0]; break;
        case 2: v = lam[2] / lam[0]; break;
        case 3: {
            double u0 = lam[0] * fabs(vec[6]) + lam[1] * fabs(vec[3]) + lam[2] * fabs(vec[0]);
            double u1 = lam[0] * fabs(vec[7]) + lam[1] * fabs(vec[4]) + lam[2] * fabs(vec[1]);
            double u2 = lam[0] * fabs(vec[8]) + lam[1] * fabs(vec[5]) + lam[2] * fabs(vec[2]);
            double nn = 0.0;
            nn += u0 * u0;
            nn += u1 * u1;
            nn += u2 * u2;
            v = u2 / sqrt(nn);
            break;
        }
        case 4: v = cbrt(lam[2] * lam[1] * lam[0]); break;
        case 5: v = (lam[0] - lam[2]) / lam[0]; break;
        case 6:
            v = -(lam[2] * log(lam[2]) + lam[1] * log(lam[1]) + lam[0] * log(lam[0]));
            break;
        case 7: v = sum; break;
        case 8: v = lam[2] / sum; break;
        case 9: v = 1 - fabs(vec[2]); break;
        }
        vals[c++] = v;
    }
    return 0;
}

} // namespace gf
