// tap_changer.cpp -- automatic tap changing over batched power flows.
//
// The reference runs its optimizer once per scenario, calling the power flow again after every tap
// change (optimizer/tap_position_optimizer.hpp:965-1110). Here every scenario keeps the optimizer's
// state between power flows, and one round runs the power flows of all scenarios as one device batch
// with each scenario's current taps; the optimizer then advances every scenario that waited for that
// batch. Scenarios that are settled are carried along at their final taps. A last batch at the final
// taps writes the caller's output, which is the output the reference's last power flow produces.
#include "tap_changer.hpp"

#include "../../include/power_grid_model_amd.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <numeric>
#include <string>
#include <thread>

namespace pgm_amd {

#define PGM_TAP_HIP_CHECK(expr)                                                                                  \
    do {                                                                                                         \
        hipError_t const e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                                    \
            throw PowerGridError(std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_) + "\n");    \
    } while (0)

namespace {

constexpr Idx kUnreached = std::numeric_limits<Idx>::max();
constexpr Idx kLastRank = kUnreached - 1;
constexpr int kMaxRounds = 100000;  // a bound every optimizer run stays far below (each rank is bounded)

std::string max_iteration_message(std::string const& what) {
    return "Maximum number of iterations reached! " + what + "\n";
}

// ---------------------------------------------------------------------------------------------------
// ranking (tap_position_optimizer.hpp:209-432): a graph over the nodes with an arc of weight 1 from the
// non-control side to the control side of every regulated transformer and arcs of weight 0 both ways
// over every other closed branch; distances from the energised sources rank the regulated transformers
// ---------------------------------------------------------------------------------------------------
struct Ranking {
    std::vector<std::vector<Idx>> ranks;  // regulator indices, rank by rank
    std::string error;                    // invalid configuration: every scenario in this state fails
};

struct TopoState {
    std::vector<IntS> line_fs, line_ts, tr_fs, tr_ts, tr_tap, src_status;
    std::vector<IntS> reg_status;
    std::vector<double> reg_u_set, reg_u_band, reg_ldc_r, reg_ldc_x;

    static TopoState of(ComponentState const& st) {
        return {st.line_fs,    st.line_ts,   st.tr_fs,      st.tr_ts,     st.tr_tap,   st.src_status,
                st.reg_status, st.reg_u_set, st.reg_u_band, st.reg_ldc_r, st.reg_ldc_x};
    }
    // the statuses that shape the ranking, as differences from `base` at the touched components
    std::string key(TopoState const& base, std::vector<std::pair<int, Idx>>& touched) const {
        std::sort(touched.begin(), touched.end());
        touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
        std::string k;
        for (auto const& [comp, i] : touched) {
            int a = 0, b = 0;
            if (comp == C_LINE) {
                a = (line_fs[i] != 0) | (line_ts[i] != 0) << 1;
                b = (base.line_fs[i] != 0) | (base.line_ts[i] != 0) << 1;
            } else if (comp == C_TRANSFORMER) {
                a = (tr_fs[i] != 0) | (tr_ts[i] != 0) << 1;
                b = (base.tr_fs[i] != 0) | (base.tr_ts[i] != 0) << 1;
            } else if (comp == C_SOURCE) {
                a = src_status[i] != 0;
                b = base.src_status[i] != 0;
            } else {
                a = reg_status[i] != 0;
                b = base.reg_status[i] != 0;
            }
            if (a != b) k += std::to_string(comp) + ':' + std::to_string(i) + '=' + std::to_string(a) + ';';
        }
        return k;
    }
    void restore(TopoState const& base, std::vector<std::pair<int, Idx>> const& touched) {
        for (auto const& [comp, i] : touched) {
            if (comp == C_LINE) {
                line_fs[i] = base.line_fs[i];
                line_ts[i] = base.line_ts[i];
            } else if (comp == C_TRANSFORMER) {
                tr_fs[i] = base.tr_fs[i];
                tr_ts[i] = base.tr_ts[i];
                tr_tap[i] = base.tr_tap[i];
            } else if (comp == C_SOURCE) {
                src_status[i] = base.src_status[i];
            } else {
                reg_status[i] = base.reg_status[i];
                reg_u_set[i] = base.reg_u_set[i];
                reg_u_band[i] = base.reg_u_band[i];
                reg_ldc_r[i] = base.reg_ldc_r[i];
                reg_ldc_x[i] = base.reg_ldc_x[i];
            }
        }
    }
};

Ranking rank_state(Model const& m, TopoState const& st) {
    Ranking out;
    Idx const n = m.n_node();
    struct Arc {
        Idx to;
        int w;
    };
    struct RegArc {
        Idx tail, head, reg;
    };
    std::vector<std::vector<Arc>> adj(static_cast<size_t>(n));
    std::vector<RegArc> reg_arcs;
    std::vector<Idx> reg_of(m.trafo.size(), -1);
    for (Idx r = 0; r < Idx(m.regulator.size()); ++r)
        if (st.reg_status[r]) reg_of[m.reg_trafo[r]] = r;
    for (Idx t = 0; t < Idx(m.trafo.size()); ++t) {
        if (!st.tr_fs[t] || !st.tr_ts[t]) continue;
        Idx const f = m.tr_f[t], to = m.tr_t[t];
        if (Idx const r = reg_of[t]; r >= 0) {
            bool const ctrl_from = m.regulator[r].control_side == 0;
            Idx const c = ctrl_from ? f : to, nc = ctrl_from ? to : f;
            adj[nc].push_back({c, 1});
            reg_arcs.push_back({nc, c, r});
        } else {
            adj[f].push_back({to, 0});
            adj[to].push_back({f, 0});
        }
    }
    // lines and links; generic branches are no edges (tap_position_optimizer.hpp:289)
    for (Idx l = 0; l < generic_slot0(m); ++l) {
        if (!st.line_fs[l] || !st.line_ts[l]) continue;
        adj[m.line_f[l]].push_back({m.line_t[l], 0});
        adj[m.line_t[l]].push_back({m.line_f[l], 0});
    }
    // shortest distance to any energised source: breadth-first over 0/1 weights
    std::vector<Idx> dist(static_cast<size_t>(n), kUnreached);
    std::deque<Idx> queue;
    for (Idx s = 0; s < Idx(m.source.size()); ++s)
        if (st.src_status[s] && dist[m.src_node[s]] != 0) {
            dist[m.src_node[s]] = 0;
            queue.push_back(m.src_node[s]);
        }
    while (!queue.empty()) {
        Idx const u = queue.front();
        queue.pop_front();
        for (Arc const& a : adj[u]) {
            Idx const d = dist[u] + a.w;
            if (d >= dist[a.to]) continue;
            dist[a.to] = d;
            if (a.w == 0)
                queue.push_front(a.to);
            else
                queue.push_back(a.to);
        }
    }
    // the reference's graph enumerates arcs by tail node
    std::stable_sort(reg_arcs.begin(), reg_arcs.end(), [](RegArc const& a, RegArc const& b) { return a.tail < b.tail; });
    struct Weighted {
        Idx weight, reg;
    };
    std::vector<Weighted> weighted;
    std::vector<ID> wrong_side;
    for (RegArc const& a : reg_arcs) {
        Idx const dt = dist[a.tail], dh = dist[a.head];
        if (dt == kUnreached && dh == kUnreached) continue;
        if (dt == kUnreached || dh == kUnreached) {
            wrong_side.push_back(m.trafo[m.reg_trafo[a.reg]].id);
            continue;
        }
        // a control side that a closer regulated transformer also reaches: lowest priority
        weighted.push_back({dt != dh - 1 ? kLastRank : dh, a.reg});
    }
    if (!wrong_side.empty()) {
        std::sort(wrong_side.begin(), wrong_side.end());
        wrong_side.erase(std::unique(wrong_side.begin(), wrong_side.end()), wrong_side.end());
        std::string msg =
            "Automatic tap changer has invalid configuration. The following transformer(s) are being controlled "
            "from non-source side towards source side:\n  Transformer IDs: ";
        for (size_t i = 0; i < wrong_side.size(); ++i) msg += (i ? ", " : "") + std::to_string(wrong_side[i]);
        out.error = msg;
        return out;
    }
    std::stable_sort(weighted.begin(), weighted.end(),
                     [](Weighted const& a, Weighted const& b) { return a.weight < b.weight; });
    Idx prev = std::numeric_limits<Idx>::min();
    for (Weighted const& w : weighted) {
        if (w.weight > prev) {
            out.ranks.emplace_back();
            prev = w.weight;
        }
        auto& r = out.ranks.back();
        if (std::find(r.begin(), r.end(), w.reg) == r.end()) r.push_back(w.reg);
    }
    return out;
}

// ---------------------------------------------------------------------------------------------------
// tap stepping (tap_position_optimizer.hpp:442-483): one step towards tap_max / tap_min, saturating;
// a higher control-side voltage needs a lower tap unless the control side is the tap side
// ---------------------------------------------------------------------------------------------------
struct TapRange {
    IntS tap_min, tap_max;
};
IntS step_up(IntS tap, TapRange r) {
    if (tap == r.tap_max) return r.tap_max;
    return IntS(r.tap_min < r.tap_max ? tap + 1 : tap - 1);
}
IntS step_down(IntS tap, TapRange r) {
    if (tap == r.tap_min) return r.tap_min;
    return IntS(r.tap_min < r.tap_max ? tap - 1 : tap + 1);
}
IntS control_voltage_up(IntS tap, TapRange r, bool control_at_tap_side) {
    return control_at_tap_side ? step_up(tap, r) : step_down(tap, r);
}
IntS control_voltage_down(IntS tap, TapRange r, bool control_at_tap_side) {
    return control_at_tap_side ? step_down(tap, r) : step_up(tap, r);
}

// voltage against the band u_set +- u_band / 2 (tap_position_optimizer.hpp:702-718): below, inside
// (also when the band is empty and the voltage sits on it) or above; unordered when NaN is involved
enum class Band { below, inside, above, unordered };
Band band_of(double v, double u_set, double u_band) {
    double const lo = u_set - 0.5 * u_band, hi = u_set + 0.5 * u_band;
    auto order = [](double a, double b) { return a < b ? -1 : a > b ? 1 : a == b ? 0 : 2; };
    int const cl = order(v, lo), ch = order(v, hi);
    if (cl != ch) return Band::inside;
    return cl == -1 ? Band::below : cl == 1 ? Band::above : cl == 0 ? Band::inside : Band::unordered;
}

// ---------------------------------------------------------------------------------------------------
// bisection over the tap range of one transformer (the reference's BinarySearch,
// tap_position_optimizer.hpp:809-944): bounds in ascending tap order, the current tap, the direction of
// the last move and two flags that end the search
// ---------------------------------------------------------------------------------------------------
class TapBisection {
  public:
    TapBisection() = default;
    TapBisection(IntS tap, TapRange r, bool control_at_tap_side) { reset(tap, r, control_at_tap_side); }

    IntS current() const { return cur_; }
    bool last_down() const { return down_; }
    bool forced() const { return forced_; }
    bool exhausted() const { return lo_ >= hi_; }
    void set_current(IntS t) { cur_ = t; }
    void clear_flags() {
        final_ = false;
        forced_ = false;
    }

    // the voltage is outside the band: move towards it (or, on the final check, to the last bound)
    void propose(bool strategy_max, bool above) {
        bool const down = (above == reversed_) != ctrl_tap_;
        if (final_) {
            cur_ = down ? lo_ : hi_;
            forced_ = true;
            return;
        }
        down_ = down;
        if (down_)
            hi_ = cur_;
        else
            lo_ = cur_;
        if (lo_ < hi_) cur_ = middle(strategy_max != reversed_);
    }
    // the voltage is inside the band: narrow the range towards the preferred end
    void narrow(bool strategy_max) {
        bool const invert = ctrl_tap_ != strategy_max;
        if (reversed_ == invert) {
            lo_ = cur_;
            down_ = false;
        } else {
            hi_ = cur_;
            down_ = true;
        }
    }
    IntS repropose(bool strategy_max, bool previous_down, bool& changed) {
        bool const prefer_higher = (strategy_max != reversed_) != ctrl_tap_;
        IntS const t = middle(prefer_higher);
        int const diff = int(t) - int(cur_);
        if (diff == 0) {
            changed = !forced_;
            forced_ = true;
            return t;
        }
        if ((diff == 1 && previous_down) || (diff == -1 && !previous_down)) final_ = true;
        changed = true;
        cur_ = t;
        return t;
    }
    void rewind(IntS tap, TapRange r) { reset(tap, r, ctrl_tap_); }

  private:
    void reset(IntS tap, TapRange r, bool control_at_tap_side) {
        down_ = false;
        final_ = false;
        forced_ = false;
        cur_ = tap;
        lo_ = std::min(r.tap_min, r.tap_max);
        hi_ = std::max(r.tap_min, r.tap_max);
        reversed_ = r.tap_max < r.tap_min;
        ctrl_tap_ = control_at_tap_side;
    }
    // midpoint of the bounds, rounded towards the preferred one (std::midpoint semantics)
    IntS middle(bool prefer_higher_in) const {
        bool const prefer_higher = ctrl_tap_ != prefer_higher_in;
        int const a = prefer_higher ? hi_ : lo_, b = prefer_higher ? lo_ : hi_;
        return IntS(a + (b - a) / 2);
    }

    IntS lo_ = 0, hi_ = 0, cur_ = 0;
    bool down_ = false, final_ = false, reversed_ = false, forced_ = false, ctrl_tap_ = false;
};

// ---------------------------------------------------------------------------------------------------
// per-scenario optimizer state
// ---------------------------------------------------------------------------------------------------
struct RegSlot {  // one ranked regulator of a scenario
    Idx reg, trafo, node;  // regulator, transformer, control-side node
    bool ctrl_from, control_at_tap_side;
    TapRange range;
    double u_set, u_band;  // per unit of the control node
    std::complex<double> z;  // line drop compensation, per unit
};

enum Method { kRequested = 0, kLinear = 1 };

struct ScenarioRun {
    int ranking = -1;
    std::vector<RegSlot> slot;           // rank by rank
    std::vector<Idx> rank_begin;          // slot offsets of the ranks (size ranks + 1)
    std::vector<uint64_t> max_range;      // largest tap range per rank
    std::vector<IntS> tap;                // current tap per slot
    std::vector<TapBisection> bis;
    std::vector<uint64_t> it_rank;        // adjusting passes per rank in the current iterate call
    int stage = 0;                        // 0 main search, 1 search around the pilot optimum
    int fallback = 0;                     // 0 first try, 1 linear method, 2 requested method again
    bool bisect = false;                  // search of the current iterate call
    int method = kRequested;              // method of the current iterate call
    int pending = -1;                     // method of the awaited power flow; -1 none
    bool failed = false;
    std::string error;
};

// voltage at the control node and current into the transformer there, from one round's results
struct RoundView {
    int P = 1;
    double base_power = 1e6;
    std::vector<int> status;
    std::vector<std::string> message;
    std::map<Idx, size_t> node_slot, trafo_slot;
    std::vector<std::vector<int8_t>> energized;
    std::vector<std::vector<double>> u_pu, u_angle, p, q;  // [slot][s * P + ph]
};

double compensated_voltage(RoundView const& rv, RegSlot const& r, Idx s) {
    size_t const ns = rv.node_slot.at(r.node), ts = rv.trafo_slot.at(r.trafo);
    double v = 0.0;
    for (int ph = 0; ph < rv.P; ++ph) {
        size_t const k = size_t(s) * rv.P + ph;
        std::complex<double> const u = std::polar(rv.u_pu[ns][k], rv.u_angle[ns][k]);
        std::complex<double> const sp(rv.p[ts][k] / rv.base_power, rv.q[ts][k] / rv.base_power);
        std::complex<double> const i = std::conj(sp / u);
        v += std::abs(u + r.z * i);
    }
    return v / rv.P;
}

// update rows of one scenario for the components that shape the ranking and the regulators
struct RowSource {
    ComponentBuffer const* buf;
    int comp;  // component slot, C_TAP_REGULATOR for regulators
    bool positional;
};

void apply_rows(Model const& m, std::vector<RowSource> const& src, Idx s, TopoState& st, std::vector<char>& rec,
                std::vector<std::pair<int, Idx>>& touched) {
    auto set_status = [](IntS& tgt, IntS v) {
        if (v != na_IntS) tgt = v != 0;
    };
    for (RowSource const& r : src) {
        rec.resize(r.buf->meta->size);
        Idx const b0 = r.buf->begin(s);
        for (Idx i = b0; i < r.buf->end(s); ++i) {
            r.buf->read_record(i, rec.data());
            ID id;
            std::memcpy(&id, rec.data(), sizeof id);
            Idx idx = r.positional ? i - b0 : find_component(m, r.comp, id);
            int comp = r.comp;
            line_slot(m, comp, idx);
            touched.emplace_back(comp, idx);
            switch (comp) {
            case C_LINE: {
                BranchUpdate u;
                std::memcpy(&u, rec.data(), sizeof u);
                set_status(st.line_fs[idx], u.from_status);
                set_status(st.line_ts[idx], u.to_status);
                break;
            }
            case C_TRANSFORMER: {
                TransformerUpdate u;
                std::memcpy(&u, rec.data(), sizeof u);
                set_status(st.tr_fs[idx], u.from_status);
                set_status(st.tr_ts[idx], u.to_status);
                if (u.tap_pos != na_IntS) st.tr_tap[idx] = tap_limit(m.tstat[idx], u.tap_pos);
                break;
            }
            case C_SOURCE: {
                SourceUpdate u;
                std::memcpy(&u, rec.data(), sizeof u);
                set_status(st.src_status[idx], u.status);
                break;
            }
            default: {
                TapRegulatorUpdate u;
                std::memcpy(&u, rec.data(), sizeof u);
                if (u.status != na_IntS) st.reg_status[idx] = u.status != 0;
                if (!std::isnan(u.u_set)) st.reg_u_set[idx] = u.u_set;
                if (!std::isnan(u.u_band)) st.reg_u_band[idx] = u.u_band;
                if (!std::isnan(u.line_drop_compensation_r)) st.reg_ldc_r[idx] = u.line_drop_compensation_r;
                if (!std::isnan(u.line_drop_compensation_x)) st.reg_ldc_x[idx] = u.line_drop_compensation_x;
            }
            }
        }
    }
}

struct TapChanger {
    Model const& m;
    PGM_Options const& opt;
    PGM_ConstDataset const* update;
    Idx S;
    bool sym;
    int P;
    double base_power;
    int strategy;  // 1 any, 2 min, 3 max, 4 fast any
    bool strategy_max, strategy_any;
    std::vector<Ranking> rankings;
    std::vector<ScenarioRun> runs;
    // device chunks handed from each round's batch to the next, and the device copies of the caller's
    // load updates that every round over all scenarios reads
    std::shared_ptr<ChunkPool> pool = std::make_shared<ChunkPool>();
    UploadCache uploads;
    // plans of the tap states met in earlier rounds
    PlanCache plan_cache;
    PlanCache* plans = &plan_cache;

    TapChanger(Model const& model, PGM_Options const& o, PGM_ConstDataset const* upd)
        : m(model), opt(o), update(upd), S(upd ? upd->batch_size : 1), sym(o.symmetric != 0), P(sym ? 1 : 3),
          base_power(sym ? kBaseP3 : kBaseP1), strategy(int(o.tap_changing_strategy)),
          strategy_max(strategy == 3), strategy_any(strategy == 1 || strategy == 4) {}

    // ---- set-up per scenario: ranking, initial taps, settings (opt_prep and pilot_run,
    // tap_position_optimizer.hpp:965-1014, 1257-1298) ----
    void prepare() {
        std::vector<RowSource> src;
        if (update) {
            for (auto const& buf : update->buffers) {
                int comp = -1;
                if (buf.meta->name == kTapRegulatorName)
                    comp = C_TAP_REGULATOR;
                else if (buf.meta->name == "line")
                    comp = C_LINE;
                else if (buf.meta->name == "link")
                    comp = C_LINK;
                else if (buf.meta->name == "transformer")
                    comp = C_TRANSFORMER;
                else if (buf.meta->name == "source")
                    comp = C_SOURCE;
                if (comp < 0) continue;
                bool all_na = buf.total_elements > 0;
                std::vector<char> rec(buf.meta->size);
                for (Idx i = 0; i < buf.total_elements && all_na; ++i) {
                    buf.read_record(i, rec.data());
                    ID id;
                    std::memcpy(&id, rec.data(), sizeof id);
                    all_na = id == na_IntID;
                }
                src.push_back({&buf, comp, all_na});
            }
        }
        TopoState const base = TopoState::of(m.base);
        std::map<std::string, int> by_key;
        std::vector<TopoState> key_state;  // one state per distinct ranking key
        runs.resize(size_t(S));
        Idx const nreg = Idx(m.regulator.size());
        // per scenario and regulator: the transformer's tap and the regulator's settings after the update
        std::vector<IntS> scn_tap(size_t(S * nreg));
        std::vector<double> scn_set(size_t(S * nreg) * 4);
        {
            TopoState work = base;
            std::vector<char> rec;
            std::vector<std::pair<int, Idx>> touched;
            for (Idx s = 0; s < S; ++s) {
                ScenarioRun& sc = runs[size_t(s)];
                touched.clear();
                try {
                    apply_rows(m, src, s, work, rec, touched);
                } catch (PowerGridError const& e) {
                    sc.failed = true;
                    sc.error = e.what();
                    work.restore(base, touched);
                    continue;
                }
                std::string const key = work.key(base, touched);
                auto it = by_key.find(key);
                if (it == by_key.end()) {
                    it = by_key.emplace(key, int(key_state.size())).first;
                    key_state.push_back(work);
                }
                sc.ranking = it->second;
                for (Idx r = 0; r < nreg; ++r) {
                    scn_tap[size_t(s * nreg + r)] = work.tr_tap[size_t(m.reg_trafo[r])];
                    double* v = scn_set.data() + size_t(s * nreg + r) * 4;
                    v[0] = work.reg_u_set[size_t(r)];
                    v[1] = work.reg_u_band[size_t(r)];
                    v[2] = work.reg_ldc_r[size_t(r)];
                    v[3] = work.reg_ldc_x[size_t(r)];
                }
                work.restore(base, touched);
            }
        }
        // rank every distinct state on several host threads
        rankings.resize(key_state.size());
        {
            unsigned const nth = unsigned(std::max<size_t>(
                1, std::min<size_t>({size_t(16), size_t(std::thread::hardware_concurrency()), key_state.size()})));
            auto work = [&](unsigned t) {
                for (size_t k = t; k < key_state.size(); k += nth) rankings[k] = rank_state(m, key_state[k]);
            };
            std::vector<std::thread> th;
            for (unsigned t = 1; t < nth; ++t) th.emplace_back(work, t);
            work(0);
            for (auto& x : th) x.join();
        }
        for (Idx s = 0; s < S; ++s) {
            ScenarioRun& sc = runs[size_t(s)];
            if (sc.failed) continue;
            Ranking const& rk = rankings[size_t(sc.ranking)];
            if (!rk.error.empty()) {
                sc.failed = true;
                sc.error = rk.error;
                continue;
            }
            sc.rank_begin.push_back(0);
            for (auto const& rank : rk.ranks) {
                uint64_t widest = 0;
                for (Idx r : rank) {
                    RegSlot slot{};
                    slot.reg = r;
                    slot.trafo = m.reg_trafo[r];
                    auto const& ts = m.tstat[slot.trafo];
                    slot.ctrl_from = m.regulator[r].control_side == 0;
                    slot.node = slot.ctrl_from ? m.tr_f[slot.trafo] : m.tr_t[slot.trafo];
                    slot.control_at_tap_side = m.regulator[r].control_side == ts.tap_side;
                    slot.range = {ts.tap_min, ts.tap_max};
                    // transformer_tap_regulator.hpp:76-88
                    double const u_rated = m.reg_u_rated[r];
                    double const* v = scn_set.data() + size_t(s * nreg + r) * 4;
                    slot.u_set = v[0] / u_rated;
                    slot.u_band = v[1] / u_rated;
                    double const z_base = u_rated * u_rated / base_power;
                    slot.z = std::complex<double>(std::isnan(v[2]) ? 0.0 : v[2], std::isnan(v[3]) ? 0.0 : v[3]) / z_base;
                    sc.slot.push_back(slot);
                    sc.tap.push_back(scn_tap[size_t(s * nreg + r)]);
                    widest = std::max<uint64_t>(
                        widest, uint64_t(std::abs(int64_t(ts.tap_max) - int64_t(ts.tap_min))));
                }
                sc.rank_begin.push_back(Idx(sc.slot.size()));
                sc.max_range.push_back(widest);
            }
            sc.bisect = strategy != 1;
            for (size_t k = 0; k < sc.slot.size(); ++k)
                sc.bis.emplace_back(sc.tap[k], sc.slot[k].range, sc.slot[k].control_at_tap_side);
            // pilot run: max / min strategies start from the extreme taps
            if (!strategy_any) {
                for (size_t k = 0; k < sc.slot.size(); ++k) {
                    RegSlot const& r = sc.slot[k];
                    bool const high_tap = strategy_max == r.control_at_tap_side;
                    sc.tap[k] = high_tap ? r.range.tap_max : r.range.tap_min;
                }
            }
            if (sc.bisect)
                for (size_t k = 0; k < sc.slot.size(); ++k) {
                    sc.bis[k].set_current(sc.tap[k]);
                    sc.bis[k].clear_flags();
                }
            start_iterate(sc, kRequested);
        }
    }

    void start_iterate(ScenarioRun& sc, int method) {
        sc.it_rank.assign(sc.max_range.size(), 0);
        sc.method = method;
        sc.pending = method;
    }
    void fail(ScenarioRun& sc, std::string const& msg) {
        sc.failed = true;
        sc.error = msg;
        sc.pending = -1;
    }

    // one regulator, linear scan (tap_position_optimizer.hpp:1132-1164)
    bool adjust_scan(ScenarioRun& sc, size_t k, RoundView const& rv, Idx s, std::vector<std::pair<size_t, IntS>>& upd) {
        RegSlot const& r = sc.slot[k];
        if (rv.energized[rv.node_slot.at(r.node)][size_t(s)] == 0) return false;
        Band const b = band_of(compensated_voltage(rv, r, s), r.u_set, r.u_band);
        IntS t = sc.tap[k];
        if (b == Band::above)
            t = control_voltage_down(sc.tap[k], r.range, r.control_at_tap_side);
        else if (b == Band::below)
            t = control_voltage_up(sc.tap[k], r.range, r.control_at_tap_side);
        if (t == sc.tap[k]) return false;
        upd.emplace_back(k, t);
        return true;
    }

    // one regulator, bisection (tap_position_optimizer.hpp:1166-1224)
    bool adjust_bisect(ScenarioRun& sc, size_t k, RoundView const& rv, Idx s,
                       std::vector<std::pair<size_t, IntS>>& upd) {
        RegSlot const& r = sc.slot[k];
        TapBisection& bs = sc.bis[k];
        if (rv.energized[rv.node_slot.at(r.node)][size_t(s)] == 0) return false;
        if (bs.exhausted() || bs.forced()) return false;
        Band const b = band_of(compensated_voltage(rv, r, s), r.u_set, r.u_band);
        if (b != Band::inside) bs.propose(strategy_max, b == Band::above);
        if (IntS const t = bs.current(); t != sc.tap[k]) {
            upd.emplace_back(k, t);
            return true;
        }
        if (strategy == 4 && b == Band::inside) return false;
        bool const previous_down = bs.last_down();
        bs.narrow(strategy_max);
        bool changed = false;
        IntS const t = bs.repropose(strategy_max, previous_down, changed);
        if (t == sc.tap[k] && b != Band::inside && !bs.exhausted()) {
            bs.rewind(t, r.range);
            throw PowerGridError(max_iteration_message(
                "TapPositionOptimizer::binary_search: no valid tap position found between tap " +
                std::to_string(int(r.range.tap_min)) + " and tap " + std::to_string(int(r.range.tap_max))));
        }
        upd.emplace_back(k, t);
        return changed;
    }

    // one pass over the ranks of TapPositionOptimizerImpl::iterate (tap_position_optimizer.hpp:1070-1108,
    // RankIteration :731-764): the first rank with a change ends the pass; true when a new power flow
    // is needed
    bool adjust_pass(ScenarioRun& sc, RoundView const& rv, Idx s) {
        std::vector<std::pair<size_t, IntS>> upd;
        bool changed = false;
        size_t const n_ranks = sc.max_range.size();
        size_t rank = 0;
        for (; rank < n_ranks; ++rank) {
            for (Idx k = sc.rank_begin[rank]; k < sc.rank_begin[rank + 1]; ++k) {
                bool const c = sc.bisect ? adjust_bisect(sc, size_t(k), rv, s, upd) : adjust_scan(sc, size_t(k), rv, s, upd);
                changed = c || changed;
            }
            if (changed) {
                std::fill(sc.it_rank.begin() + Idx(rank) + 1, sc.it_rank.end(), 0);
                ++sc.it_rank[rank];
                break;
            }
        }
        if (!changed) return false;
        if (sc.it_rank[rank] > 2 * sc.max_range[rank])
            throw PowerGridError(max_iteration_message(
                "TapPositionOptimizer::iterate " + std::to_string(sc.it_rank[rank]) + " iterations reached: " +
                std::to_string(sc.max_range[rank]) + "x2 iterations in rank " + std::to_string(rank)));
        for (auto const& [k, t] : upd) sc.tap[k] = tap_limit(m.tstat[sc.slot[k].trafo], t);
        return true;
    }

    // the awaited power flow of a scenario is in (iterate_with_fallback and optimize,
    // tap_position_optimizer.hpp:1016-1068): adjust, or move on to the next iterate call, or finish
    void advance(ScenarioRun& sc, RoundView const& rv, Idx s) {
        int const st = rv.status[size_t(s)];
        if (st == 3) {  // the scenario's update data itself is invalid
            fail(sc, rv.message[size_t(s)]);
            return;
        }
        std::string problem;
        bool again = false;
        if (st != 0) {
            problem = rv.message[size_t(s)];
        } else {
            try {
                again = adjust_pass(sc, rv, s);
            } catch (PowerGridError const& e) {
                problem = e.what();
            }
        }
        if (!problem.empty()) {
            // diverging or singular power flows and the iteration limit fall back to one pass with the
            // linear method followed by the requested method again; a second failure is final
            if (sc.fallback == 0) {
                sc.fallback = 1;
                start_iterate(sc, kLinear);
            } else {
                fail(sc, problem);
            }
            return;
        }
        if (again) {
            sc.pending = sc.method;
            return;
        }
        if (sc.fallback == 1) {
            sc.fallback = 2;
            start_iterate(sc, kRequested);
            return;
        }
        if (sc.stage == 0 && !strategy_any) {
            // exploit_neighborhood (tap_position_optimizer.hpp:1300-1328): one step further towards the
            // preferred end, then a linear scan back into the band
            for (size_t k = 0; k < sc.slot.size(); ++k) {
                RegSlot const& r = sc.slot[k];
                IntS const t = strategy_max ? control_voltage_up(sc.tap[k], r.range, r.control_at_tap_side)
                                            : control_voltage_down(sc.tap[k], r.range, r.control_at_tap_side);
                sc.tap[k] = tap_limit(m.tstat[r.trafo], t);
            }
            sc.stage = 1;
            sc.fallback = 0;
            sc.bisect = false;
            start_iterate(sc, kRequested);
            return;
        }
        sc.pending = -1;
    }

    // the update data of one round over the scenarios `scns` (global indices, ascending): the caller's
    // buffers (less the regulators) followed by the taps of every scenario's regulated transformers, which
    // therefore take precedence. Over every scenario the caller's buffers are passed on as they are;
    // over a subset their rows are copied, scenario by scenario, into row-based buffers.
    struct RoundData {
        PGM_ConstDataset ds;
        std::vector<TransformerUpdate> rows;
        std::vector<Idx> indptr;
        std::vector<std::vector<char>> copies;
        std::vector<std::vector<Idx>> copy_indptr;
    };
    void build_round(RoundData& rd, std::vector<Idx> const& scns) const {
        Idx const k = Idx(scns.size());
        bool const all = k == S;
        rd.ds.meta = &meta_dataset("update");
        rd.ds.is_batch = true;
        rd.ds.batch_size = k;
        if (update) {
            rd.copies.reserve(update->buffers.size());
            rd.copy_indptr.reserve(update->buffers.size());
            for (auto const& buf : update->buffers) {
                if (buf.meta->name == kTapRegulatorName) continue;
                if (all) {
                    rd.ds.buffers.push_back(buf);
                    continue;
                }
                size_t const rs = buf.meta->size;
                Idx n = 0;
                for (Idx s : scns) n += buf.end(s) - buf.begin(s);
                auto& data = rd.copies.emplace_back(size_t(n) * rs);
                auto& ind = rd.copy_indptr.emplace_back(1, 0);
                Idx row = 0;
                for (Idx s : scns) {
                    if (buf.columnar()) {
                        for (Idx i = buf.begin(s); i < buf.end(s); ++i, ++row)
                            buf.read_record(i, data.data() + size_t(row) * rs);
                    } else {
                        size_t const cnt = size_t(buf.end(s) - buf.begin(s));
                        std::memcpy(data.data() + size_t(row) * rs,
                                    static_cast<char const*>(buf.data) + size_t(buf.begin(s)) * rs, cnt * rs);
                        row += Idx(cnt);
                    }
                    ind.push_back(row);
                }
                ComponentBuffer cb;
                cb.meta = buf.meta;
                cb.data = data.data();
                cb.total_elements = n;
                if (buf.elements_per_scenario >= 0) {
                    cb.elements_per_scenario = buf.elements_per_scenario;
                } else {
                    cb.elements_per_scenario = -1;
                    cb.indptr = ind.data();
                }
                rd.ds.buffers.push_back(cb);
            }
        }
        rd.rows.clear();
        rd.indptr.assign(1, 0);
        bool uniform = true;
        for (Idx j = 0; j < k; ++j) {
            ScenarioRun const& sc = runs[size_t(scns[size_t(j)])];
            for (size_t q = 0; q < sc.slot.size(); ++q)
                rd.rows.push_back(TransformerUpdate{m.trafo[sc.slot[q].trafo].id, na_IntS, na_IntS, sc.tap[q]});
            rd.indptr.push_back(Idx(rd.rows.size()));
            uniform = uniform && rd.indptr[size_t(j) + 1] - rd.indptr[size_t(j)] == rd.indptr[1];
        }
        if (rd.rows.empty()) return;
        ComponentBuffer tb;
        tb.meta = rd.ds.meta->find("transformer");
        tb.data = rd.rows.data();
        tb.total_elements = Idx(rd.rows.size());
        if (uniform) {
            tb.elements_per_scenario = rd.indptr[1];
        } else {
            tb.elements_per_scenario = -1;
            tb.indptr = rd.indptr.data();
        }
        rd.ds.buffers.push_back(tb);
    }
    std::vector<Idx> every_scenario() const {
        std::vector<Idx> v(static_cast<size_t>(S));
        std::iota(v.begin(), v.end(), Idx(0));
        return v;
    }
    PGM_Options round_options(int method) const {
        PGM_Options o = opt;
        o.tap_changing_strategy = 0;
        if (method == kLinear) o.calculation_method = PGM_linear;
        return o;
    }

    // one power-flow batch over the scenarios `scns` at their current taps; the control-node voltages
    // and the control-side transformer flows are gathered for the optimizer, indexed by position in scns
    void run_round(int method, std::vector<Idx> const& scns, RoundView& rv) {
        auto const t0 = std::chrono::steady_clock::now();
        RoundData rd;
        build_round(rd, scns);
        Idx const S = Idx(scns.size());
        // only the caller's own buffers (rounds over every scenario) may go through the upload cache: the
        // copies of a subset round are temporaries whose addresses can recur with other contents
        BatchContext ctx;
        ctx.pool = pool;
        ctx.uploads = S == this->S ? &uploads : nullptr;
        ctx.plans = plans;
        auto b = create_batch(m, round_options(method), &rd.ds, ctx);
        auto const t1 = std::chrono::steady_clock::now();
        PGM_MutableDataset out;
        out.meta = &meta_dataset(sym ? "sym_output" : "asym_output");
        out.is_batch = true;
        out.batch_size = S;
        auto columns = [&](char const* comp, std::vector<char const*> names) {
            ComponentBuffer cb;
            cb.meta = out.meta->find(comp);
            cb.elements_per_scenario = m.count(std::string(comp) == "node" ? C_NODE : C_TRANSFORMER);
            cb.total_elements = cb.elements_per_scenario * S;
            for (char const* n : names) cb.columns.emplace_back(cb.meta->find(n), nullptr);
            out.buffers.push_back(cb);
        };
        columns("node", {"energized", "u_pu", "u_angle"});
        columns("transformer", {"p_from", "q_from", "p_to", "q_to"});
        b->set_output(out);
        b->run();
        rv = RoundView{};
        rv.P = P;
        rv.base_power = base_power;
        rv.status = b->status;
        rv.message = b->messages;
        for (Idx s : scns)
            for (auto const& r : runs[size_t(s)].slot) {
                rv.node_slot.emplace(r.node, rv.node_slot.size());
                rv.trafo_slot.emplace(r.trafo, rv.trafo_slot.size());
            }
        // every needed (column, item) is packed on the device by one strided copy each, then the packed
        // block comes back in one transfer
        struct Pick {
            void const* src;
            size_t size, pitch;
            void* dst;
        };
        std::vector<Pick> picks;
        auto pick = [&](int comp, char const* name, Idx item, void* dst) {
            for (auto const& dc : b->cols[comp]) {
                if (dc.attr->name != name) continue;
                size_t const as = dc.attr->size;
                picks.push_back({static_cast<char const*>(dc.dev) + size_t(item) * as, as,
                                 size_t(m.count(comp)) * as, dst});
                return;
            }
            throw PowerGridError(std::string("missing output column ") + name);
        };
        rv.energized.resize(rv.node_slot.size());
        rv.u_pu.resize(rv.node_slot.size());
        rv.u_angle.resize(rv.node_slot.size());
        for (auto const& [node, k] : rv.node_slot) {
            rv.energized[k].resize(size_t(S));
            rv.u_pu[k].resize(size_t(S) * P);
            rv.u_angle[k].resize(size_t(S) * P);
            pick(C_NODE, "energized", node, rv.energized[k].data());
            pick(C_NODE, "u_pu", node, rv.u_pu[k].data());
            pick(C_NODE, "u_angle", node, rv.u_angle[k].data());
        }
        rv.p.resize(rv.trafo_slot.size());
        rv.q.resize(rv.trafo_slot.size());
        for (auto const& [t, k] : rv.trafo_slot) {
            bool from = false;
            for (Idx r = 0; r < Idx(m.regulator.size()); ++r)
                if (m.reg_trafo[r] == t) from = m.regulator[r].control_side == 0;
            rv.p[k].resize(size_t(S) * P);
            rv.q[k].resize(size_t(S) * P);
            pick(C_TRANSFORMER, from ? "p_from" : "p_to", t, rv.p[k].data());
            pick(C_TRANSFORMER, from ? "q_from" : "q_to", t, rv.q[k].data());
        }
        size_t packed = 0;
        for (auto const& pk : picks) packed += pk.size * size_t(S);
        if (packed > 0) {
            void* dev = nullptr;
            PGM_TAP_HIP_CHECK(hipMalloc(&dev, packed));
            std::vector<char> host(packed);
            try {
                size_t off = 0;
                for (auto const& pk : picks) {
                    PGM_TAP_HIP_CHECK(hipMemcpy2D(static_cast<char*>(dev) + off, pk.size, pk.src, pk.pitch, pk.size,
                                                  size_t(S), hipMemcpyDeviceToDevice));
                    off += pk.size * size_t(S);
                }
                PGM_TAP_HIP_CHECK(hipMemcpy(host.data(), dev, packed, hipMemcpyDeviceToHost));
            } catch (...) {
                (void)hipFree(dev);
                throw;
            }
            PGM_TAP_HIP_CHECK(hipFree(dev));
            size_t off = 0;
            for (auto const& pk : picks) {
                std::memcpy(pk.dst, host.data() + off, pk.size * size_t(S));
                off += pk.size * size_t(S);
            }
        }
        // a scenario that failed to power-flow leaves its columns unwritten: mark it not energised
        for (Idx s = 0; s < S; ++s)
            if (rv.status[size_t(s)] != 0)
                for (auto& e : rv.energized) e[size_t(s)] = 0;
        if (std::getenv("PGM_AMD_TRACE")) {
            auto const t2 = std::chrono::steady_clock::now();
            std::fprintf(stderr, "[pgm_amd] tap round (method %d, %lld scenarios): batch %8.1f ms, run + gather %8.1f ms\n",
                         method, (long long)S, std::chrono::duration<double, std::milli>(t1 - t0).count(),
                         std::chrono::duration<double, std::milli>(t2 - t1).count());
        }
    }

    void optimize() {
        prepare();
        for (int round = 0;; ++round) {
            bool any = false;
            for (int method : {int(kRequested), int(kLinear)}) {
                std::vector<Idx> waiting;
                for (Idx s = 0; s < S; ++s)
                    if (runs[size_t(s)].pending == method) waiting.push_back(s);
                if (waiting.empty()) continue;
                any = true;
                if (round >= kMaxRounds) {
                    for (Idx s : waiting) fail(runs[size_t(s)], max_iteration_message("TapPositionOptimizer::rounds"));
                    continue;
                }
                // a minority of waiting scenarios runs alone; otherwise every scenario runs (the caller's
                // buffers need no copy then)
                std::vector<Idx> const scns = 2 * waiting.size() <= size_t(S) ? waiting : every_scenario();
                RoundView rv;
                run_round(method, scns, rv);
                for (size_t j = 0, w = 0; j < scns.size() && w < waiting.size(); ++j)
                    if (scns[j] == waiting[w]) {
                        advance(runs[size_t(waiting[w])], rv, Idx(j));
                        ++w;
                    }
            }
            if (!any) break;
        }
    }

    // the caller's output at the final taps, the failures of the optimizer and the regulators' taps
    void finish(PGM_MutableDataset const& output, PGM_Handle* handle) {
        RoundData rd;
        build_round(rd, every_scenario());
        BatchContext ctx;
        ctx.pool = pool;
        ctx.uploads = &uploads;
        ctx.plans = plans;
        auto b = create_batch(m, round_options(kRequested), &rd.ds, ctx);
        b->is_batch = update != nullptr;
        b->set_output(output);
        b->run();
        // a scenario the optimizer gave up on reports the optimizer's error (whatever the last batch did)
        for (Idx s = 0; s < S; ++s) {
            ScenarioRun const& sc = runs[size_t(s)];
            if (sc.failed) {
                b->status[size_t(s)] = 4;
                b->messages[size_t(s)] = sc.error;
            }
        }
        b->fetch(output, handle);
        ComponentBuffer const* rb = output.find(kTapRegulatorName);
        if (!rb || m.regulator.empty()) return;
        Idx const n = Idx(m.regulator.size());
        for (Idx s = 0; s < S; ++s) {
            if (b->status[size_t(s)] != 0) continue;
            ScenarioRun const& sc = runs[size_t(s)];
            if (rb->end(s) - rb->begin(s) != n)
                throw PowerGridError("output buffer of transformer_tap_regulator must hold every element per scenario\n");
            for (Idx r = 0; r < n; ++r) {
                TapRegulatorOutput o{m.regulator[r].id, 0, na_IntS};
                for (size_t k = 0; k < sc.slot.size(); ++k)
                    if (sc.slot[k].reg == r) {
                        o.energized = 1;
                        o.tap_pos = sc.tap[k];
                    }
                rb->write_record(rb->begin(s) + r, &o);
            }
        }
    }
};

}  // namespace

void calculate_with_tap_changer(Model const& model, PGM_Options const& opt, PGM_MutableDataset const& output,
                                PGM_ConstDataset const* update, PGM_Handle* handle) {
    if (opt.tap_changing_strategy < 1 || opt.tap_changing_strategy > 4)
        throw PowerGridError("get_optimizer_type is not implemented for " + std::to_string(opt.tap_changing_strategy) +
                             "!\n");
    if (opt.calculation_type != PGM_power_flow)
        throw PowerGridError("automatic tap changing is defined for power flow only\n");
    TapChanger tc(model, opt, update);
    tc.optimize();
    tc.finish(output, handle);
}

void write_null_fault_output(Model const& m, PGM_MutableDataset const& output) {
    ComponentBuffer const* fb = output.find(kFaultName);
    if (!fb) return;
    Idx const n = Idx(m.fault.size());
    for (Idx s = 0; s < output.batch_size; ++s)
        for (Idx i = fb->begin(s); i < fb->end(s) && i - fb->begin(s) < n; ++i) {
            FaultOutput const o{m.fault[i - fb->begin(s)].id, 0};
            fb->write_record(i, &o);
        }
}

void write_null_regulator_output(Model const& m, PGM_MutableDataset const& output) {
    ComponentBuffer const* rb = output.find(kTapRegulatorName);
    if (!rb) return;
    Idx const n = Idx(m.regulator.size());
    for (Idx s = 0; s < output.batch_size; ++s)
        for (Idx i = rb->begin(s); i < rb->end(s) && i - rb->begin(s) < n; ++i) {
            TapRegulatorOutput const o{m.regulator[i - rb->begin(s)].id, 0, na_IntS};
            rb->write_record(i, &o);
        }
}

std::vector<Idx> tap_regulator_ranks(Model const& model, ComponentState const& st) {
    Ranking const rk = rank_state(model, TopoState::of(st));
    if (!rk.error.empty()) throw PowerGridError(rk.error);
    std::vector<Idx> out(model.regulator.size(), -1);
    for (size_t i = 0; i < rk.ranks.size(); ++i)
        for (Idx r : rk.ranks[i]) out[size_t(r)] = Idx(i);
    return out;
}

}  // namespace pgm_amd
