// Host planner of the ia/ibk series approximations (frm_iaibk.hpp): the
// structural steps of calc_va (fr_d_t_g_2d.cpp:490-741, constant steps,
// blocks of one step) on slot numbers instead of coefficient arrays.  Every
// decision below (validity ranges from ia_mini / ibk_i2, splits, series of
// new centres, merges, removals) is the reference's, in its order; the
// coefficient arithmetic becomes operations the GPU applies per node.
#include "frm_iaibk.hpp"

#include <algorithm>
#include <cmath>

namespace frm {
namespace {

constexpr double kApproxEps = 1e-7;  // approx_eps (252)
constexpr double kFixEps = 1e-09;    // eps of the centre iterations (253)

// ia_mini (256-275): x beyond which the nterms-th ia term is below approx_eps
double ia_mini(double A, double r, int nterms, double a, double b) {
    double ak = 1;
    for (int i = 1; i < nterms; i++) ak *= (A - i + 1) / i;
    return std::pow(std::fabs(ak * (std::pow(a - r, (double)nterms) - std::pow(b - r, (double)nterms)) /
                              kApproxEps),
                    1.0 / (nterms - A));
}

// ibk_i2 (352-384, constant steps): half-width of an ibk's validity
double ibk_i2(double A, double r, double bc, int nterms, double a, double c) {
    double ak = 1;
    for (int i = 1; i < nterms; i++) ak *= (A - i + 1) / i;
    if (std::fabs(bc - r) < 1e-10) return std::pow(std::fabs(kApproxEps / ak), 1.0 / nterms);
    const double p1 = std::pow((a + bc - r), A) / std::pow((a + bc - r), nterms);
    const double p2 = std::pow((c + bc - r), A) / std::pow((c + bc - r), nterms);
    return std::pow(std::fabs(kApproxEps / (ak * (p1 - p2))), 1.0 / nterms);
}

struct Ia {
    double min;
    int nterms, slot;
};
struct Ibk {
    double min, max, b;
    int nterms, old, slot;
};

struct Planner {
    double A, a, b;
    IaPlan& P;
    std::vector<Ia> ias;
    std::vector<Ibk> ibks;
    std::vector<int> free_slots;
    int used = 0;

    int alloc() {
        int s;
        if (!free_slots.empty()) {
            s = free_slots.back();
            free_slots.pop_back();
        } else {
            s = used++;
        }
        P.nslots = std::max(P.nslots, used);
        return s;
    }
    void release(int s) { free_slots.push_back(s); }
    void emit(int kind, int slot, int src, int nt, int off) { P.ops.push_back({kind, slot, src, nt, off}); }
    int push_vals(const double* v, int n) {
        const int off = (int)P.vals.size();
        P.vals.insert(P.vals.end(), v, v + n);
        return off;
    }
    // factors of an ibk's coefficients for block r (its krk / F): the
    // difference of the two powers around centre bc (407-421)
    int ibk_factors(double bc, int r, int nt) {
        double d[kIaMaxAk];
        for (int i = 0; i < nt; i++)
            d[i] = std::pow(a + bc - (double)r, A - (double)i) - std::pow(b + bc - (double)r, A - (double)i);
        return push_vals(d, nt);
    }
    // factors of an ia's coefficients for block r (296-307); c[0] = 0
    int ia_factors(int r, int nt) {
        double d[kIaMaxAk];
        d[0] = 0.0;
        for (int i = 1; i < nt; i++)
            d[i] = std::pow(a - (double)r, (double)i) - std::pow(b - (double)r, (double)i);
        return push_vals(d, nt);
    }
    double ibk_min(double bc, double rmax, int nt) { return bc - std::floor(ibk_i2(A, rmax, bc, nt, a, b)); }
    // the fixed point bb = rr + ibk_i2(r, bb) + 1 that sets the next centre
    double next_centre(double rr, double r) {
        double bb0 = rr, bb = rr + 1;
        do {
            bb0 = bb;
            bb = rr + ibk_i2(A, r, bb, kIaAk, a, b) + 1;
        } while ((bb - bb0) > kFixEps);
        if (!std::isfinite(bb)) {
            bb = bb0;
            if (!std::isfinite(bb)) bb = rr;
        }
        return bb;
    }
    // a new ibk from block r around centre floor(bb) with minimum rr
    Ibk new_ibk(double cb, double rr, int r) {
        Ibk nb{rr, cb, cb, kIaAk, 0, alloc()};
        emit(kIaCreate, nb.slot, 0, kIaAk, ibk_factors(cb, r, kIaAk));
        return nb;
    }

    // calc_va(A, 1, block, ...): evaluate at x = block + 1, then fold the
    // increment of block + 1 into the approximations
    void step(int block) {
        const double rmax = block + 2;
        const int rmaxidx = (int)rmax;
        const int r = block + 1;
        const double x = block + 1;
        // evaluation: ias valid at x, then ibks whose range holds x (519-526)
        for (const Ia& e : ias)
            if (x >= e.min) {
                double v[kIaMaxAk];
                for (int i = 0; i < e.nterms; i++) v[i] = std::pow(x, A - (double)i);
                emit(kIaEvalIa, e.slot, 0, e.nterms, push_vals(v, e.nterms));
            }
        for (const Ibk& e : ibks)
            if ((x >= e.min) && (x <= e.max)) {
                double v[kIaMaxAk], xbk = 1;
                for (int i = 0; i < e.nterms; i++) {
                    v[i] = xbk;
                    xbk *= (x - e.b);
                }
                emit(kIaEvalIbk, e.slot, 0, e.nterms, push_vals(v, e.nterms));
            }
        {
            const double w = std::pow((double)r - (double)r + a, A) - std::pow((double)r - (double)r + b, A);
            emit(kIaExact, 0, 0, 1, push_vals(&w, 1));
        }
        // 1) drop the ibks that end before rmax
        for (size_t j = 0; j < ibks.size(); j++)
            if (ibks[j].max < rmax) {
                release(ibks[j].slot);
                ibks.erase(ibks.begin() + j);
                j--;
            }
        // 2) the block joins the ibks beyond rmax; split those it cannot
        // cover to approx_eps
        int ss = (int)ibks.size();
        for (int j = 0; j < ss; j++)
            if (ibks[j].min >= rmax && ibks[j].old == 0) {
                int spl = 0, bnt = kIaAk;
                double bmin = ibk_min(ibks[j].b, rmax, kIaAk);
                Ibk i2{};
                if (bmin > ibks[j].min) {
                    int aa;
                    for (aa = kIaAk; aa < kIaMaxAk; aa++)
                        if ((ibks[j].b - ibk_i2(A, rmaxidx, ibks[j].b, aa, a, b)) < ibks[j].min) break;
                    if (aa == kIaMaxAk) {
                        spl = 1;
                    } else {
                        bnt = aa;
                        bmin = ibk_min(ibks[j].b, rmax, aa);
                    }
                    if (spl) {
                        i2 = ibks[j];
                        i2.slot = alloc();
                        emit(kIaCopy, i2.slot, ibks[j].slot, kIaMaxAk, 0);
                    }
                }
                emit(kIaAdd, ibks[j].slot, 0, bnt, ibk_factors(ibks[j].b, r, bnt));
                if (ibks[j].nterms < bnt) ibks[j].nterms = bnt;
                if (spl) {
                    i2.max = std::ceil(bmin);
                    ibks[j].min = i2.max + 1;
                    i2.old = 1;
                    if (ibks[j].min > ibks[j].max) {
                        i2.max = ibks[j].max;
                        release(ibks[j].slot);
                        ibks.erase(ibks.begin() + j);
                        j--;
                        ss--;
                    }
                    ibks.push_back(i2);
                    double rr = i2.min, bb;
                    do {
                        bb = next_centre(rr, i2.min);
                        Ibk nb = new_ibk(std::floor(bb), rr, r);
                        rr = std::floor(bb) + 1;
                        if (std::floor(bb) > i2.max) nb.max = i2.max;
                        ibks.push_back(nb);
                    } while (std::floor(bb) < i2.max);
                }
            }
        // 3) a series of new ibks for the range of the ibk straddling rmax
        double bm1 = -1;
        for (Ibk& e : ibks)
            if (e.max >= rmax && e.min < rmax && e.max > bm1 && e.old == 0) {
                bm1 = e.max;
                e.old = 1;
            }
        if (bm1 != -1) {
            for (Ibk& e : ibks)
                if (e.max < bm1) e.old = 1;
            double rr = rmax, bb;
            do {
                bb = next_centre(rr, rmaxidx);
                Ibk nb = new_ibk(std::floor(bb), rr, r);
                rr = std::floor(bb) + 1;
                if (std::floor(bb) > bm1) nb.max = bm1;
                ibks.push_back(nb);
            } while (std::floor(bb) < bm1);
        }
        // 4) the block's ia joins the last ia if valid from earlier on
        double bia_min = std::ceil(ia_mini(A, rmax, kIaAk, a, b));
        if (!std::isfinite(bia_min)) bia_min = rmax;
        if (!ias.empty() && bia_min < ias.back().min) {
            emit(kIaAdd, ias.back().slot, 0, kIaAk, ia_factors(r, kIaAk));
            if (ias.back().nterms < kIaAk) ias.back().nterms = kIaAk;
            return;
        }
        // 5) otherwise ibks bridge up to the new ia's minimum
        bm1 = rmax;
        for (const Ibk& e : ibks)
            if ((e.max + 1) > bm1) bm1 = e.max + 1;
        double rr = bm1, bb;
        do {
            bb = next_centre(rr, rmaxidx);
            ibks.push_back(new_ibk(std::floor(bb), rr, r));
            rr = std::floor(bb) + 1;
        } while (std::floor(bb) < std::ceil(bia_min));
        Ia bia{rr, kIaAk, alloc()};
        emit(kIaCreate, bia.slot, 0, kIaAk, ia_factors(r, kIaAk));
        ias.push_back(bia);
    }
};

}  // namespace

IaPlan iaibk_plan(double A, double a, double b, int steps) {
    IaPlan P;
    double ak = 1;
    P.ak[0] = 1;
    for (int i = 1; i < kIaMaxAk; i++) {
        ak *= (A - i + 1) / i;
        P.ak[i] = ak;
    }
    Planner pl{A, a, b, P};
    P.step_off.push_back(0);
    for (int n = 1; n <= steps; ++n) {
        pl.step(n - 1);
        P.step_off.push_back((int)P.ops.size());
    }
    return P;
}

}  // namespace frm
