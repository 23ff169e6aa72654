// ed_tridiag.h -- eigenpairs of a real symmetric tridiagonal matrix by the
// implicit QL algorithm with Wilkinson shifts (Bowdler, Martin, Reinsch and
// Wilkinson, Numer. Math. 11 (1968) 293).  Host code; m is the Lanczos
// iteration count, so O(m^2) without vectors and O(m^3) with them.
#pragma once
#include <cmath>
#include <vector>
#include <algorithm>

namespace edc {

// d: diagonal [m] (overwritten by ascending eigenvalues); e: off-diagonal
// [m-1] (destroyed).  If z is non-null it receives the eigenvectors,
// column-major z[i + k*m] = component i of eigenvector k.  Returns false if
// an eigenvalue needed more than 60 sweeps.
inline bool tridiag_ql(std::vector<double>& d, std::vector<double> e, std::vector<double>* z) {
    const int m = (int)d.size();
    if (m == 0) return true;
    e.resize(m, 0.0);
    e[m - 1] = 0.0;
    if (z) {
        z->assign((size_t)m * m, 0.0);
        for (int i = 0; i < m; i++) (*z)[i + (size_t)i * m] = 1.0;
    }
    for (int l = 0; l < m; l++) {
        int iter = 0;
        while (true) {
            int mm = l;
            for (; mm < m - 1; mm++) {
                double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= 1e-16 * dd) break;
            }
            if (mm == l) break;
            if (++iter > 60) return false;
            // Wilkinson shift from the leading 2x2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = mm - 1;
            bool deflated = false;
            for (; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (z) {
                    double* zi = z->data() + (size_t)i * m;
                    double* zi1 = z->data() + (size_t)(i + 1) * m;
                    for (int k = 0; k < m; k++) {
                        double t = zi1[k];
                        zi1[k] = s * zi[k] + c * t;
                        zi[k] = c * zi[k] - s * t;
                    }
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    // ascending order
    std::vector<int> idx(m);
    for (int i = 0; i < m; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    std::vector<double> ds(m);
    for (int i = 0; i < m; i++) ds[i] = d[idx[i]];
    if (z) {
        std::vector<double> zs((size_t)m * m);
        for (int k = 0; k < m; k++)
            for (int i = 0; i < m; i++) zs[i + (size_t)k * m] = (*z)[i + (size_t)idx[k] * m];
        *z = zs;
    }
    d = ds;
    return true;
}

// Eigenpairs of a dense real symmetric matrix A (m x m, row-major, read
// symmetric from its upper triangle) by the cyclic Jacobi method with the
// threshold strategy of Rutishauser (Handbook for Automatic Computation II,
// contribution II/1): rotations zero the off-diagonal entries sweep by sweep
// until their sum is below the rounding level.  d receives the eigenvalues
// ascending, z (column-major, z[i + k*m]) the eigenvectors.  For the
// thick-restart Lanczos matrices (arrowhead + tridiagonal, m up to a few
// hundred).  Returns false if 100 sweeps did not converge.
inline bool symeig_jacobi(std::vector<double> A, int m, std::vector<double>& d, std::vector<double>& z) {
    d.assign(m, 0.0);
    z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) {
        z[i + (size_t)i * m] = 1.0;
        d[i] = A[(size_t)i * m + i];
        for (int j = 0; j < i; j++) A[(size_t)i * m + j] = A[(size_t)j * m + i];
    }
    std::vector<double> b(d), zz(m, 0.0);
    bool ok = false;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) off += std::fabs(A[(size_t)p * m + q]);
        if (off == 0.0) { ok = true; break; }
        const double tresh = sweep < 3 ? 0.2 * off / ((double)m * m) : 0.0;
        for (int p = 0; p < m; p++) {
            for (int q = p + 1; q < m; q++) {
                double& apq = A[(size_t)p * m + q];
                const double g = 100.0 * std::fabs(apq);
                if (sweep > 3 && std::fabs(d[p]) + g == std::fabs(d[p]) && std::fabs(d[q]) + g == std::fabs(d[q])) {
                    apq = 0.0;
                    continue;
                }
                if (std::fabs(apq) <= tresh) continue;
                double h = d[q] - d[p];
                double t;
                if (std::fabs(h) + g == std::fabs(h)) {
                    t = apq / h;
                } else {
                    const double theta = 0.5 * h / apq;
                    t = 1.0 / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
                    if (theta < 0.0) t = -t;
                }
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c, tau = s / (1.0 + c);
                h = t * apq;
                zz[p] -= h; zz[q] += h;
                d[p] -= h; d[q] += h;
                apq = 0.0;
                auto rot = [&](double& x, double& y) {
                    const double gx = x, hy = y;
                    x = gx - s * (hy + gx * tau);
                    y = hy + s * (gx - hy * tau);
                };
                for (int j = 0; j < p; j++) rot(A[(size_t)j * m + p], A[(size_t)j * m + q]);
                for (int j = p + 1; j < q; j++) rot(A[(size_t)p * m + j], A[(size_t)j * m + q]);
                for (int j = q + 1; j < m; j++) rot(A[(size_t)p * m + j], A[(size_t)q * m + j]);
                for (int j = 0; j < m; j++) rot(z[j + (size_t)p * m], z[j + (size_t)q * m]);
            }
        }
        for (int p = 0; p < m; p++) {
            b[p] += zz[p];
            d[p] = b[p];
            zz[p] = 0.0;
        }
    }
    // ascending order
    std::vector<int> idx(m);
    for (int i = 0; i < m; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return d[a] < d[c]; });
    std::vector<double> d2(m), z2((size_t)m * m);
    for (int k = 0; k < m; k++) {
        d2[k] = d[idx[k]];
        for (int i = 0; i < m; i++) z2[i + (size_t)k * m] = z[i + (size_t)idx[k] * m];
    }
    d.swap(d2);
    z.swap(z2);
    return ok;
}

}  // namespace edc
