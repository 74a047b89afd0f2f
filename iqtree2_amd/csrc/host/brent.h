// Optimization::minimizeOneDimen and brent_opt (utils/optimization.cpp:
// 185-373) as a resumable state machine: the caller evaluates f at x() and
// feeds the value back, so the function (one tree likelihood per call, on
// the device) stays outside. Brent's method is the published one (Brent
// 1973, the Numerical Recipes form the reference uses: golden-section steps
// of CGOLD with parabolic interpolation, ITMAX 100, ZEPS 1e-10); the
// bracketing around the guess (eps = 50 * tol * guess, falling back to the
// whole interval when the guess is not lower than both neighbours) and the
// final "keep the guess if Brent ended higher" rule are the reference's.
#pragma once
#include <cmath>

namespace plk {

class BrentRun {
public:
    // minimizeOneDimen(xmin, xguess, xmax, tolerance)
    void start(double xmin, double xguess, double xmax, double tol) {
        xmin_ = xmin;
        xmax_ = xmax;
        tol_ = tol;
        if (xguess < xmin) xguess = xmin;
        if (xguess > xmax) xguess = xmax;
        const double eps = xguess * tol * 50.0;
        ax_ = xguess - eps;
        if (ax_ < xmin) ax_ = xmin;
        bx_ = xguess;
        cx_ = xguess + eps;
        if (cx_ > xmax) cx_ = xmax;
        evals_ = 0;
        done_ = widen_ = false;
        ask(S_FB, bx_);
    }
    bool done() const { return done_; }
    double x() const { return next_; }        // abscissa of the next evaluation (while !done())
    double result() const { return result_; }
    double fx() const { return fresult_; }    // f at result()
    double f2x() const { return f2_; }        // brent_opt's parabolic second derivative estimate
    int evals() const { return evals_; }

    void feed(double f) {
        evals_++;
        switch (state_) {
        case S_FB:
            fb_ = f;
            if (ax_ < bx_) return ask(S_FA, ax_);
            fa_ = fb_;
            return after_fa();
        case S_FA:
            fa_ = f;
            return after_fa();
        case S_FC:
            fc_ = f;
            return after_fc();
        case S_FMIN:
            fa_ = f;
            return widen_max();
        case S_FMAX:
            fc_ = f;
            return brent_start();
        case S_FU:
            return brent_update(f);
        case S_FBX:
            // Brent ended above f(guess): the guess, re-evaluated
            result_ = bx_;
            fresult_ = f;
            done_ = true;
            return;
        }
    }

private:
    enum State { S_FB, S_FA, S_FC, S_FMIN, S_FMAX, S_FU, S_FBX };
    static constexpr int kItmax = 100;
    static constexpr double kCgold = 0.3819660;
    static constexpr double kZeps = 1.0e-10;
    static double sign(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

    void ask(State s, double x) {
        state_ = s;
        next_ = x;
    }
    void after_fa() {
        if (bx_ < cx_) return ask(S_FC, cx_);
        fc_ = fb_;
        after_fc();
    }
    void after_fc() {
        if (fa_ < fb_ || fc_ < fb_) {
            widen_ = true;   // the guess is not a bracket: the whole interval
            if (ax_ != xmin_) return ask(S_FMIN, xmin_);
            return widen_max();
        }
        brent_start();
    }
    void widen_max() {
        if (cx_ != xmax_) return ask(S_FMAX, xmax_);
        brent_start();
    }
    // brent_opt(ax, bx, cx, tol, ..., fa, fb, fc) after the bracket
    void brent_start() {
        if (widen_) {
            ax_ = xmin_;
            cx_ = xmax_;
        }
        a_ = ax_ < cx_ ? ax_ : cx_;
        b_ = ax_ > cx_ ? ax_ : cx_;
        x_ = bx_;
        fxb_ = fb_;
        if (fa_ < fc_) {
            w_ = ax_;
            fw_ = fa_;
            v_ = cx_;
            fv_ = fc_;
        } else {
            w_ = cx_;
            fw_ = fc_;
            v_ = ax_;
            fv_ = fa_;
        }
        d_ = 0.0;
        e_ = 0.0;
        iter_ = 1;
        brent_step();
    }
    void brent_finish() {
        const double xw = x_ - w_, wv = w_ - v_, vx = v_ - x_;
        f2_ = 2.0 * (fv_ * xw + fxb_ * wv + fw_ * vx) / (v_ * v_ * xw + x_ * x_ * wv + w_ * w_ * vx);
        // minimizeOneDimen: keep the guess when Brent ended higher
        if (fxb_ > fb_) return ask(S_FBX, bx_);
        result_ = x_;
        fresult_ = fxb_;
        done_ = true;
    }
    void brent_step() {
        if (iter_ > kItmax) return brent_finish();
        const double xm = 0.5 * (a_ + b_);
        tol1_ = tol_ * std::fabs(x_) + kZeps;
        const double tol2 = 2.0 * tol1_;
        if (std::fabs(x_ - xm) <= (tol2 - 0.5 * (b_ - a_))) return brent_finish();
        if (std::fabs(e_) > tol1_) {
            double r = (x_ - w_) * (fxb_ - fv_);
            double q = (x_ - v_) * (fxb_ - fw_);
            double p = (x_ - v_) * q - (x_ - w_) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p;
            q = std::fabs(q);
            const double etemp = e_;
            e_ = d_;
            if (std::fabs(p) >= std::fabs(0.5 * q * etemp) || p <= q * (a_ - x_) || p >= q * (b_ - x_)) {
                d_ = kCgold * (e_ = (x_ >= xm ? a_ - x_ : b_ - x_));
            } else {
                d_ = p / q;
                const double u = x_ + d_;
                if (u - a_ < tol2 || b_ - u < tol2) d_ = sign(tol1_, xm - x_);
            }
        } else {
            d_ = kCgold * (e_ = (x_ >= xm ? a_ - x_ : b_ - x_));
        }
        u_ = std::fabs(d_) >= tol1_ ? x_ + d_ : x_ + sign(tol1_, d_);
        ask(S_FU, u_);
    }
    void brent_update(double fu) {
        if (fu <= fxb_) {
            if (u_ >= x_) a_ = x_;
            else b_ = x_;
            v_ = w_;
            w_ = x_;
            x_ = u_;
            fv_ = fw_;
            fw_ = fxb_;
            fxb_ = fu;
        } else {
            if (u_ < x_) a_ = u_;
            else b_ = u_;
            if (fu <= fw_ || w_ == x_) {
                v_ = w_;
                w_ = u_;
                fv_ = fw_;
                fw_ = fu;
            } else if (fu <= fv_ || v_ == x_ || v_ == w_) {
                v_ = u_;
                fv_ = fu;
            }
        }
        iter_++;
        brent_step();
    }

    State state_ = S_FB;
    double xmin_ = 0, xmax_ = 0, tol_ = 0, next_ = 0;
    double ax_ = 0, bx_ = 0, cx_ = 0, fa_ = 0, fb_ = 0, fc_ = 0;
    double a_ = 0, b_ = 0, x_ = 0, w_ = 0, v_ = 0, u_ = 0, d_ = 0, e_ = 0, tol1_ = 0;
    double fxb_ = 0, fw_ = 0, fv_ = 0, f2_ = 0;
    double result_ = 0, fresult_ = 0;
    int iter_ = 0, evals_ = 0;
    bool done_ = false, widen_ = false;
};

}  // namespace plk
