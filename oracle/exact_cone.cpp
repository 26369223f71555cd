// TEST INFRASTRUCTURE — FP64 exact-cone contact oracle.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load or run anything under oracle/; the product never does.
//
// Solves the convex cone program the reference uses as its exact-Coulomb
// baseline (src/core/ReferenceConicSolver.cpp:182-430): impulses lambda, per
// contact (normal, tangent u, tangent v), minimise
//     1/2 lambda^T H lambda + l^T lambda,
//     H = J M^-1 J^T + R (R diagonal, positive),  l = J v_free - target,
// subject to lambda_c in the friction cone of contact c, then reports the
// post-impulse velocity v = v_free + M^-1 J^T lambda. Elliptic cones
// ||(lambda_u / mu_u, lambda_v / mu_v)|| <= lambda_n are solved in scaled
// variables s = (lambda_n, lambda_u / mu_u, lambda_v / mu_v), in which the cone
// is the unit second-order cone and the program stays convex (H becomes
// D H D with D = diag(1, mu_u, mu_v)). A zero friction coefficient fixes that
// tangent impulse at zero.
//
// Method: projected gradient with Nesterov/FISTA extrapolation (Beck and
// Teboulle, 2009), step 1/L with L the largest absolute row sum of the scaled
// Hessian (an upper bound of its spectral norm), restarted from the last
// iterate whenever the objective increases. Stopping rule: the gradient
// mapping L * || s - P(s - grad / L) || divided by (1 + ||l||) at or below the
// tolerance. Projection onto the unit second-order cone: the closed form of
// the Euclidean projection (inside: unchanged; inside the polar cone: zero;
// else ((n + ||t||) / 2) * (1, t / ||t||)).
//
// Everything is double precision.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Program {
    uint32_t n = 0;  // 3 * contacts
    std::vector<double> H, l, scale;  // scaled Hessian, linear term, D
    std::vector<uint8_t> fixed;       // 1: component fixed at zero (zero friction)
};

double dotv(const double* a, const double* b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

void project_unit_cone(double* s, const uint8_t* fixed) {
    if (fixed[1]) s[1] = 0.0;
    if (fixed[2]) s[2] = 0.0;
    const double t = std::hypot(s[1], s[2]);
    const double nn = s[0];
    if (t <= nn) return;  // inside (also covers t == 0, n >= 0)
    if (t <= -nn) {       // polar cone
        s[0] = s[1] = s[2] = 0.0;
        return;
    }
    const double a = 0.5 * (nn + t);
    s[0] = a;
    s[1] *= a / t;
    s[2] *= a / t;
}

void project_all(std::vector<double>& s, const Program& p) {
    for (uint32_t c = 0; 3u * c < p.n; ++c) project_unit_cone(&s[3u * c], &p.fixed[3u * c]);
}

void hess_mul(const Program& p, const std::vector<double>& x, std::vector<double>& y) {
    for (uint32_t r = 0; r < p.n; ++r) y[r] = dotv(&p.H[(size_t)r * p.n], x.data(), p.n);
}

double objective(const Program& p, const std::vector<double>& x, std::vector<double>& tmp) {
    hess_mul(p, x, tmp);
    return 0.5 * dotv(x.data(), tmp.data(), p.n) + dotv(p.l.data(), x.data(), p.n);
}

} // namespace

namespace {

// Solves the scaled program; fills impulses (unscaled) and stats.
int solve_program(Program& p, uint32_t contacts, const double* warm, uint32_t max_iterations, double tolerance,
                  double* impulses, double* stats) {
    const uint32_t n = p.n;
    double L = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
        double rs = 0.0;
        for (uint32_t c = 0; c < n; ++c) rs += std::fabs(p.H[(size_t)r * n + c]);
        L = std::max(L, rs);
    }
    if (!(L > 0.0) || !std::isfinite(L)) return 10;
    const double step = 1.0 / L;
    const double rscale = 1.0 + std::sqrt(dotv(p.l.data(), p.l.data(), n));

    std::vector<double> cur(n, 0.0), ext(n), cand(n), grad(n), tmp(n), pstep(n);
    double used_warm = 0.0;
    if (warm != nullptr) {
        std::vector<double> w(n);
        for (uint32_t r = 0; r < n; ++r) w[r] = p.scale[r] > 0.0 ? warm[r] / p.scale[r] : 0.0;
        project_all(w, p);
        if (objective(p, w, tmp) < 0.0) {
            cur = w;
            used_warm = 1.0;
        }
    }
    ext = cur;
    double t = 1.0;
    double fcur = objective(p, cur, tmp);
    int code = 11;
    double residual = INFINITY;
    uint32_t it = 0;
    for (; it < max_iterations; ++it) {
        hess_mul(p, ext, grad);
        for (uint32_t r = 0; r < n; ++r) cand[r] = ext[r] - step * (grad[r] + p.l[r]);
        project_all(cand, p);
        double fc = objective(p, cand, tmp);
        if (fc > fcur + 1.0e-14) {  // restart from the last iterate
            ext = cur;
            t = 1.0;
            hess_mul(p, ext, grad);
            for (uint32_t r = 0; r < n; ++r) cand[r] = ext[r] - step * (grad[r] + p.l[r]);
            project_all(cand, p);
            fc = objective(p, cand, tmp);
        }
        // gradient mapping at the candidate
        hess_mul(p, cand, grad);
        for (uint32_t r = 0; r < n; ++r) pstep[r] = cand[r] - step * (grad[r] + p.l[r]);
        project_all(pstep, p);
        double gm = 0.0;
        for (uint32_t r = 0; r < n; ++r) {
            const double d = L * (cand[r] - pstep[r]);
            gm += d * d;
        }
        residual = std::sqrt(gm) / rscale;
        if (!std::isfinite(residual) || !std::isfinite(fc)) return 3;
        const double tn = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
        for (uint32_t r = 0; r < n; ++r) ext[r] = cand[r] + ((t - 1.0) / tn) * (cand[r] - cur[r]);
        t = tn;
        cur = cand;
        fcur = fc;
        if (residual <= tolerance) {
            code = 0;
            ++it;
            break;
        }
    }
    // cone violation in impulse units: max(-lambda_n, mu (||s_t|| - lambda_n)),
    // which for a circular cone is the usual max(-lambda_n, ||lambda_t|| - mu lambda_n)
    double violation = 0.0;
    for (uint32_t c = 0; c < contacts; ++c) {
        const double* s = &cur[3u * c];
        const double mu = std::max(p.scale[3u * c + 1u], p.scale[3u * c + 2u]);
        violation = std::max(violation, std::max(-s[0], mu * (std::hypot(s[1], s[2]) - s[0])));
    }
    for (uint32_t r = 0; r < n; ++r) impulses[r] = cur[r] * p.scale[r];
    stats[0] = (double)it;
    stats[1] = residual;
    stats[2] = violation;
    stats[3] = used_warm;
    stats[4] = fcur;
    return code;
}

int set_cones(Program& p, uint32_t contacts, const double* friction_uv) {
    for (uint32_t c = 0; c < contacts; ++c) {
        const double mu[2] = {friction_uv[2u * c], friction_uv[2u * c + 1u]};
        for (int k = 0; k < 2; ++k) {
            if (!(mu[k] >= 0.0) || !std::isfinite(mu[k])) return 2;
            p.scale[3u * c + 1u + k] = mu[k] > 0.0 ? mu[k] : 0.0;
            p.fixed[3u * c + 1u + k] = mu[k] > 0.0 ? 0u : 1u;
        }
    }
    return 0;
}

} // namespace

extern "C" {

// The program given directly: H (n x n, symmetric positive definite incl.
// regularisation) and l (n), n = 3 * contacts. Same codes and stats as below.
int mr_oracle_exact_cone_program(uint32_t contacts, const double* hessian, const double* linear,
                                 const double* friction_uv, const double* warm, uint32_t max_iterations,
                                 double tolerance, double* impulses, double* stats) {
    const uint32_t n = 3u * contacts;
    Program p;
    p.n = n;
    p.H.assign((size_t)n * n, 0.0);
    p.l.assign(n, 0.0);
    p.scale.assign(n, 1.0);
    p.fixed.assign(n, 0u);
    if (set_cones(p, contacts, friction_uv) != 0) return 2;
    for (uint32_t r = 0; r < n; ++r) {
        if (!std::isfinite(linear[r])) return 2;
        p.l[r] = linear[r] * p.scale[r];
        for (uint32_t c = 0; c < n; ++c) {
            if (!std::isfinite(hessian[(size_t)r * n + c])) return 2;
            p.H[(size_t)r * n + c] = p.scale[r] * hessian[(size_t)r * n + c] * p.scale[c];
        }
    }
    return solve_program(p, contacts, warm, max_iterations, tolerance, impulses, stats);
}

// Returns 0 on convergence, 11 (did not converge) when the iteration budget
// ends first, 10 when H is not usable, 2 on non-finite input.
// stats: [iterations, scaled optimality residual, max cone violation,
//         warm start used (0/1), objective].
int mr_oracle_exact_cone(uint32_t nv, const double* inverse_mass, const double* free_velocity, uint32_t contacts,
                         const double* jacobian, const double* target, const double* regularization,
                         const double* friction_uv, const double* warm, uint32_t max_iterations, double tolerance,
                         double* impulses, double* velocity, double* stats) {
    const uint32_t n = 3u * contacts;
    for (uint32_t i = 0; i < nv * nv; ++i)
        if (!std::isfinite(inverse_mass[i])) return 2;
    for (uint32_t i = 0; i < n * nv; ++i)
        if (!std::isfinite(jacobian[i])) return 2;
    // M^-1 J^T (nv x n), H = J M^-1 J^T + R, l = J v_free - target
    std::vector<double> MJt((size_t)nv * n, 0.0);
    for (uint32_t r = 0; r < n; ++r)
        for (uint32_t i = 0; i < nv; ++i) {
            double s = 0.0;
            for (uint32_t k = 0; k < nv; ++k) s += inverse_mass[(size_t)i * nv + k] * jacobian[(size_t)r * nv + k];
            MJt[(size_t)i * n + r] = s;
        }
    std::vector<double> H((size_t)n * n), l(n);
    for (uint32_t r = 0; r < n; ++r) {
        double jv = 0.0;
        for (uint32_t k = 0; k < nv; ++k) jv += jacobian[(size_t)r * nv + k] * free_velocity[k];
        l[r] = jv - target[r];
        for (uint32_t c = 0; c < n; ++c) {
            double h = 0.0;
            for (uint32_t k = 0; k < nv; ++k) h += jacobian[(size_t)r * nv + k] * MJt[(size_t)k * n + c];
            if (r == c) h += regularization[r];
            H[(size_t)r * n + c] = h;
        }
    }
    const int code = mr_oracle_exact_cone_program(contacts, H.data(), l.data(), friction_uv, warm, max_iterations,
                                                  tolerance, impulses, stats);
    if (code == 2 || code == 3 || code == 10) return code;
    for (uint32_t i = 0; i < nv; ++i) {
        double dv = 0.0;
        for (uint32_t r = 0; r < n; ++r) dv += MJt[(size_t)i * n + r] * impulses[r];
        velocity[i] = free_velocity[i] + dv;
    }
    return code;
}

}  // extern "C"
