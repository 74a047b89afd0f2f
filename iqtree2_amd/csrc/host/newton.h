// Optimization::minimizeNewton (utils/optimization.cpp:422-499) as a
// resumable state machine: the caller evaluates f = -dlnL/dt and
// f' = -d2lnL/dt2 at x() and feeds them back, so many branch optimisations
// can advance one derivative evaluation at a time in a shared batch. The
// arithmetic and control flow are those of the reference routine: bracket
// [x1, x2], Newton step unless it leaves the bracket or f' <= 0 (then
// bisection), return rts_old when |dx| < xacc or at the last iteration.
#pragma once
#include <cmath>

namespace plk {

class NewtonRun {
public:
    // rtsafe(x1 = lo, x_init, x2 = hi, xacc, maxit)
    void start(double x_init, double lo, double hi, double xacc, int max_nr) {
        x1_ = lo;
        x2_ = hi;
        xacc_ = xacc;
        max_nr_ = max_nr;
        rts_ = x_init;
        if (rts_ < x1_) rts_ = x1_;
        if (rts_ > x2_) rts_ = x2_;
        first_ = true;
        done_ = failed_ = false;
        evals_ = 0;
    }
    bool done() const { return done_; }
    // maximum number of iterations reached without a result (max_nr < 1)
    bool failed() const { return failed_; }
    double result() const { return result_; }
    int evals() const { return evals_; }
    // abscissa of the next derivative evaluation (valid while !done())
    double x() const { return rts_; }
    // abscissa of the last derivative evaluation: the branch length that
    // PhyloTree::computeFuncDerv left in the tree, which the divergence guard
    // of optimizeOneBranch evaluates (tree/phylotree.cpp:2616-2618). It equals
    // result() except when a bisection step makes no progress.
    double last_x() const { return last_x_; }

    void feed(double f, double df) {
        evals_++;
        last_x_ = rts_;
        f_ = f;
        df_ = df;
        if (first_) {
            first_ = false;
            if (df_ >= 0.0 && std::fabs(f_) < xacc_) return finish(rts_);
            if (f_ < 0.0) {
                xl_ = rts_;
                xh_ = x2_;
            } else {
                xh_ = rts_;
                xl_ = x1_;
            }
            dx_ = dxold_ = std::fabs(xh_ - xl_);
            j_ = 1;
            if (j_ > max_nr_) {
                failed_ = done_ = true;
                return;
            }
            return advance();
        }
        if (df_ > 0.0 && std::fabs(f_) < xacc_) return finish(rts_);
        if (f_ < 0.0) xl_ = rts_;
        else if (f_ > 0.0) xh_ = rts_;
        j_++;
        advance();
    }

private:
    void finish(double x) {
        result_ = x;
        done_ = true;
    }
    // body of the iteration loop up to the next function evaluation
    void advance() {
        rts_old_ = rts_;
        if ((df_ <= 0.0) || (((rts_ - xh_) * df_ - f_) * ((rts_ - xl_) * df_ - f_) >= 0.0)) {
            dxold_ = dx_;
            dx_ = 0.5 * (xh_ - xl_);
            rts_ = xl_ + dx_;
            if (xl_ == rts_) return finish(rts_);
        } else {
            dxold_ = dx_;
            dx_ = f_ / df_;
            const double temp = rts_;
            rts_ -= dx_;
            if (temp == rts_) return finish(rts_);
        }
        if (std::fabs(dx_) < xacc_ || j_ == max_nr_) return finish(rts_old_);
    }

    double x1_ = 0, x2_ = 0, xacc_ = 0, rts_ = 0, rts_old_ = 0, xl_ = 0, xh_ = 0;
    double dx_ = 0, dxold_ = 0, f_ = 0, df_ = 0, result_ = 0, last_x_ = 0;
    int j_ = 0, max_nr_ = 0, evals_ = 0;
    bool first_ = true, done_ = false, failed_ = false;
};

}  // namespace plk
