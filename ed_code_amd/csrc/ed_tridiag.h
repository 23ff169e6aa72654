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

}  // namespace edc
