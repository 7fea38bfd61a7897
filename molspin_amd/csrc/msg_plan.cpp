// Execution-plan resolution for the device path: the property list of an MSD
// Task block ("key=value; key=value; ...") -> msg_plan.
//
// Restates ResolveExecutionPlan (RunSection/General/HS/HSExecutionPlan.cpp:89-615)
// for the keys of the static Hilbert-space path, with the reference's defaults
// (HSExecutionPlan.h:33-117), aliases, value spellings and error messages.  Keys
// that select physics outside this path (powder grids, secular approximations,
// spin polarisation / CIDSP output, pulse sequences)
// are reported as MSG_UNSUPPORTED with the key named, so a caller can fall back
// to its CPU path.  Property splitting follows ObjectParser
// (MSDParser/ObjectParser.cpp:37-86): ';' ends a value unless inside ( ).
#include <cctype>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/molspin_gpu.h"
#include "msg_props.hpp"

namespace {

using msg::lower;
using msg::one_of;
using msg::Props;
using msg::trim;

struct PlanError {
    msg_status st;
    std::string msg;
};

void resolve(const Props &p, msg_plan &out) {
    auto invalid = [](const std::string &m) { return PlanError{MSG_INVALID_ARG, m}; };
    auto unsupported = [](const std::string &m) { return PlanError{MSG_UNSUPPORTED, m}; };
    msg_run_params &r = out.params;
    std::memset(&out, 0, sizeof(out));
    // defaults (HSExecutionPlan.h:33-117)
    r.calculation = MSG_CALC_TIMEEVOLUTION;
    r.observables = MSG_OBS_STATE_POPULATIONS;
    r.method = MSG_METHOD_EXPONENTIAL;
    r.krylov_size = 16;
    r.sampling = MSG_SAMPLING_DIRECT;
    r.yield_corrections = 0;
    r.total_time = 10000.0;
    r.time_step = 1.0;
    r.seed = 1;
    r.task = MSG_TASK_HSGENERAL;
    out.propagation = MSG_PROP_EXPONENTIAL;
    out.montecarlo_samples = 1;
    out.sampling_method = MSG_TRACE_SUZ;
    out.auto_seed = 1;
    out.device_backend = 0;
    out.time_infinity = 0;

    std::string v;
    // task type (RunSection_CreateTask.cpp:82-88,122-125,242-253)
    if (p.str({"type"}, v)) {
        if (v == "hsgeneral") r.task = MSG_TASK_HSGENERAL;
        else if (v == "statichs-direct-yields") r.task = MSG_TASK_LEGACY_DIRECT_YIELDS;
        else if (v == "statichs-stoch-yields") r.task = MSG_TASK_LEGACY_STOCH_YIELDS;
        else if (v == "statichs-symmetricdecay" || v == "statichs") r.task = MSG_TASK_SYMMETRIC_DECAY;
        else throw unsupported("task type " + v + " is not on the device path");
    }
    if (p.str({"backend"}, v)) {
        if (v == "mi355x") out.device_backend = 1;
        else if (v != "cpu") throw invalid("backend must be cpu or mi355x");
    }
    if (r.task != MSG_TASK_HSGENERAL) {
        // the legacy StaticHS tasks compute yields on their own grids
        r.calculation = MSG_CALC_YIELDS;
        r.observables = MSG_OBS_TRANSITION_YIELDS;
        r.method = MSG_METHOD_KRYLOV;
        out.propagation = MSG_PROP_KRYLOV;
        p.num("totaltime", r.total_time);
        p.num("timestep", r.time_step);
        bool b = false;
        if (p.boolean("yieldcorrections", b)) r.yield_corrections = b;
        if (p.str({"propagationmethod"}, v)) {
            if (v == "krylov") { r.method = MSG_METHOD_KRYLOV; out.propagation = MSG_PROP_KRYLOV; }
            else if (v == "autoexpm") { r.method = MSG_METHOD_EXPONENTIAL; out.propagation = MSG_PROP_AUTOEXPM; }
            else if (one_of(v, {"normal", "exp", "exponential"})) { r.method = MSG_METHOD_EXPONENTIAL; out.propagation = MSG_PROP_EXPONENTIAL; }
            else throw unsupported("propagationmethod " + v + " is not on the device path");
        }
        int ks = 16;
        if (p.integer("krylovsize", ks) && ks >= 1) r.krylov_size = ks;
        // krylovtol (TaskStaticHSDirectYields.cpp:373-410): the Krylov basis is kept while
        // h_{m+1,m} |c_{m-1}| <= krylovtol; unset or not positive, the reference's 1e-16
        double kt = 0.0;
        if (p.num("krylovtol", kt) && std::isfinite(kt) && kt > 0.0) r.krylov_tol = kt;
        int mc = 1;
        if (p.integer("montecarlosamples", mc)) {
            if (mc < 1) throw invalid("montecarlosamples must be at least one");
            out.montecarlo_samples = mc;
        }
        if (p.str({"samplingmethod"}, v)) {
            if (v == "coherent") out.sampling_method = MSG_TRACE_COHERENT;
            else if (v != "suz") throw invalid("samplingmethod must be suz or coherent");
        }
        if (r.task == MSG_TASK_LEGACY_STOCH_YIELDS) r.sampling = MSG_SAMPLING_GIVEN;
        if (p.boolean("autoseed", b)) out.auto_seed = b;
        double seed = 1.0;
        if (p.num("seed", seed)) r.seed = (uint64_t)std::llround(std::fabs(seed));
        if (!(std::isfinite(r.total_time) && r.total_time >= 0.0)) throw invalid("totaltime must be finite and non-negative");
        if (!(std::isfinite(r.time_step) && r.time_step > 0.0)) throw invalid("timestep must be finite and positive");
        if (r.krylov_size > 32) throw unsupported("krylovsize above 32 is not on the device path");
        return;
    }

    // HSGeneral (HSExecutionPlan.cpp:95-615).  Inputs the reference rejects are
    // rejected here with its message (MSG_INVALID_ARG); valid plans that select
    // physics off the device path are reported once validation has finished.
    std::string offpath;
    auto off = [&](const std::string &why) {
        if (offpath.empty()) offpath = why;
    };
    const bool has_h0 = p.has("hamiltonianh0list");
    v = "static";
    p.str({"dynamics", "timedependence", "time_dependence"}, v);
    if (one_of(v, {"timeindependent", "time-independent"})) v = "static";
    if (one_of(v, {"timedependent", "time-dependent"})) v = "dynamic";
    const bool dynamic = v == "dynamic";
    if (!dynamic && v != "static") throw invalid("dynamics must be static or dynamic");
    out.params.dynamics = dynamic ? 1 : 0;

    v = "timeevolution";
    p.str({"calculation", "calculationmode", "calculation_mode"}, v);
    if (one_of(v, {"timeevo", "time-evolution", "time_evolution"})) v = "timeevolution";
    if (one_of(v, {"yield", "quantumyield", "quantumyields"})) v = "yields";
    if (one_of(v, {"spectra", "spectrum", "resonance"}))
        throw invalid("spectroscopy is a standalone task; use StaticHS-Direct-Spectra (or StaticHS-Resonance-Spectra "
                      "for the dedicated resonance algorithm)");
    if (v == "timeevolution") r.calculation = MSG_CALC_TIMEEVOLUTION;
    else if (v == "yields") r.calculation = MSG_CALC_YIELDS;
    else throw invalid("calculation must be timeevolution or yields");

    v = "direct";
    p.str({"sampling", "tracesampling", "trace_sampling"}, v);
    if (one_of(v, {"trace", "montecarlo", "monte-carlo", "mc"})) v = "stochastic";
    if (v == "direct") r.sampling = MSG_SAMPLING_DIRECT;
    else if (v == "stochastic") r.sampling = MSG_SAMPLING_GIVEN;  // columns drawn on the host
    else throw invalid("sampling must be direct or stochastic");

    std::string approx = "full";
    p.str({"approximation", "hamiltonianapproximation", "hamiltonian_approximation"}, approx);
    bool secular = false;
    if (p.boolean("secularization", secular) || p.boolean("secular", secular)) approx = secular ? "secular" : "full";
    if (one_of(approx, {"rwa", "rotatingwave", "rotating-wave"}))
        throw invalid("approximation=rwa is not a static Hamiltonian secular approximation; configure the "
                      "oscillatory drive Interaction/Pulse explicitly and use approximation=full or secular/highfield");
    if (one_of(approx, {"highfield", "high-field"})) approx = "secular";
    if (one_of(approx, {"nonsecular", "non-secular", "exact"})) approx = "full";
    if (approx != "full" && approx != "secular") throw invalid("approximation must be full or secular");
    if (approx == "secular") {
        if (!has_h0)
            throw invalid("approximation=secular/RWA requires hamiltonianh0list so H0 is explicit; hamiltonianh1list "
                          "is optional and remains non-secular");
        off("approximation=secular is not on the device path");
    }

    // powder sampling: the reference's simple consistency rules, then off path
    std::string grid;
    const bool grid_given = p.str({"powdergrid", "powder_grid", "powdergridtype"}, grid);
    if (grid_given && !one_of(grid, {"uniform", "golden", "fibonacci", "sophe", "octant"}))
        throw invalid("powdergrid must be uniform, sophe, or octant");
    const bool generated = grid_given || p.has("powdersamplingpoints") || p.has("powdergridsize");
    int gamma = 1;
    if (p.integer("powdergammapoints", gamma) && gamma > 1 && !generated)
        throw invalid("powdergammapoints greater than one requires a generated theta/phi powder grid");
    const bool explicit_orientation = p.has("powderorientation") || p.has("orientation");
    bool powder = false;
    const bool powder_given = p.boolean("powderaveraging", powder) || p.boolean("powder", powder);
    if (powder_given && powder && !generated && !explicit_orientation)
        throw invalid("powderaveraging=true requires a generated powder grid or an explicit powderorientation");
    if (generated || explicit_orientation) {
        if (powder_given && !powder)
            throw invalid("powderaveraging=false conflicts with explicit powder sampling keywords");
        if (!has_h0)
            throw invalid("HSGeneral powder propagation requires hamiltonianh0list so the orientation-specific H0 "
                          "is explicit");
        off("powder averaging is not on the device path");
    }

    p.num("totaltime", r.total_time);
    p.num("timestep", r.time_step);
    if (!(std::isfinite(r.total_time) && r.total_time >= 0.0)) throw invalid("totaltime must be finite and non-negative");
    if (!(std::isfinite(r.time_step) && r.time_step > 0.0)) throw invalid("timestep must be finite and positive");

    v = "normal";
    p.str({"propagationmethod", "propagator"}, v);
    if (one_of(v, {"normal", "exp", "exponential", "matrixexponential", "matrix-exponential"})) {
        r.method = MSG_METHOD_EXPONENTIAL;
        out.propagation = MSG_PROP_EXPONENTIAL;
    } else if (v == "autoexpm") {
        r.method = MSG_METHOD_EXPONENTIAL;
        out.propagation = MSG_PROP_AUTOEXPM;
    } else if (v == "krylov") {
        r.method = MSG_METHOD_KRYLOV;
        out.propagation = MSG_PROP_KRYLOV;
    } else if (one_of(v, {"rk4", "explicit"})) {
        r.method = MSG_METHOD_RK4;
        out.propagation = MSG_PROP_RK4;
    } else if (v == "uncoupled") {  // this backend's exact uncoupled-radical propagator
        r.method = MSG_METHOD_UNCOUPLED;
        out.propagation = MSG_PROP_EXPONENTIAL;
    } else {
        throw invalid("propagationmethod must be normal/exp, autoexpm, krylov, or rk4");
    }
    int ks = 16;
    if (p.integer("krylovsize", ks)) r.krylov_size = ks < 1 ? 16 : ks;
    if (r.krylov_size > 32) off("krylovsize above 32 is not on the device path");
    double ktol = 0.0;
    if (p.has("krylovtol") && !(p.num("krylovtol", ktol) && std::isfinite(ktol) && ktol > 0.0))
        throw invalid("krylovtol must be finite and positive when supplied");

    int mc = 1;
    if (p.integer("montecarlosamples", mc) && mc < 1) throw invalid("montecarlosamples must be at least one");
    out.montecarlo_samples = mc;
    v = "suz";
    p.str({"samplingmethod"}, v);
    if (v == "coherent") out.sampling_method = MSG_TRACE_COHERENT;
    else if (v != "suz") throw invalid("samplingmethod must be suz or coherent");
    bool b = true;
    if (p.boolean("autoseed", b)) out.auto_seed = b;
    double seed = 1.0;
    if (p.num("seed", seed)) r.seed = (uint64_t)std::llround(std::fabs(seed));

    for (const char *k : {"hamiltonianh0list", "hamiltonianh1list", "initialstatehamiltonian", "pulsesequence"})
        if (p.has(k)) off(std::string(k) + " is not on the device path");
    const bool spinlist = p.has("spinlist");
    bool transition_yields = true;
    const bool ty_given = p.boolean("transitionyields", transition_yields);
    if (p.boolean("yieldcorrections", b)) r.yield_corrections = b;
    bool cidsp = false, cidnp = false;
    const bool cidsp_given = p.boolean("cidsp", cidsp);
    const bool cidnp_given = p.boolean("cidnp", cidnp);
    if (cidnp_given && cidnp) cidsp = true;

    std::string mode = "auto";
    if (p.str({"observables", "observablemode", "observable_mode"}, mode)) {
        if (one_of(mode, {"states", "state", "populations", "statepopulations", "state-populations"})) mode = "states";
        else if (one_of(mode, {"spins", "spin", "polarization", "polarizations", "spinpolarization", "spin-polarization"}))
            mode = "spins";
        else if (one_of(mode, {"transitionyields", "transition-yields", "yields", "quantumyields", "quantum-yields"}))
            mode = "transitionyields";
        else if (one_of(mode, {"cidsp", "cidnp", "productpolarization", "product-polarization"})) mode = "cidsp";
        else if (mode != "auto") throw invalid("observables must be auto, states, spins, transitionyields, or cidsp");
    }
    bool polarization = false;  // spin-polarisation or CIDSP observables
    if (mode == "auto") {
        if (!ty_given && spinlist) transition_yields = false;
        polarization = spinlist;
    } else if (mode == "states") {
        if (r.calculation != MSG_CALC_TIMEEVOLUTION) throw invalid("observables=states requires calculation=timeevolution");
        if (spinlist || cidsp) throw invalid("observables=states cannot be combined with spinlist or cidsp/cidnp");
        if (ty_given && transition_yields) throw invalid("observables=states conflicts with transitionyields=true");
        transition_yields = false;
    } else if (mode == "spins") {
        if (!spinlist) throw invalid("observables=spins requires spinlist");
        if ((cidsp_given && cidsp) || (cidnp_given && cidnp)) throw invalid("observables=spins conflicts with cidsp/cidnp=true");
        if (ty_given && transition_yields) throw invalid("observables=spins conflicts with transitionyields=true");
        transition_yields = false;
        cidsp = false;
        polarization = true;
    } else if (mode == "transitionyields") {
        if (r.calculation != MSG_CALC_YIELDS) throw invalid("observables=transitionyields requires calculation=yields");
        if (spinlist || cidsp)
            throw invalid("observables=transitionyields cannot be combined with spinlist or cidsp/cidnp");
        if (ty_given && !transition_yields)
            throw invalid("observables=transitionyields conflicts with transitionyields=false");
        transition_yields = true;
    } else {
        if (!spinlist) throw invalid("observables=cidsp requires spinlist");
        if (ty_given && transition_yields) throw invalid("observables=cidsp conflicts with transitionyields=true");
        transition_yields = false;
        cidsp = true;
        polarization = true;
    }
    if (cidsp && !spinlist) throw invalid("cidsp/cidnp output requires spinlist");
    if (r.calculation == MSG_CALC_YIELDS && !transition_yields && !spinlist)
        throw invalid("calculation=yields with transitionyields=false requires spinlist for integrated "
                      "polarization/CIDSP output");
    if (polarization) off("spin-polarisation / CIDSP observables are not on the device path");
    // observables that remain (HSObservableCollector.cpp:84-99)
    r.observables = (mode == "transitionyields" || (mode == "auto" && r.calculation == MSG_CALC_YIELDS && transition_yields))
                        ? MSG_OBS_TRANSITION_YIELDS
                        : MSG_OBS_STATE_POPULATIONS;

    std::string tw;
    for (const char *k : {"printtimeframe", "integrationtimeframe"})
        if (p.str({k}, tw)) {
            if (one_of(tw, {"pulse", "pulses", "preparation"}))
                throw invalid("printtimeframe=pulse requires a task-level pulsesequence");
            if (!one_of(tw, {"freeevo", "free-evolution", "free_evolution", "free", "full", "all"}))
                throw invalid(std::string(k) + " must be pulse, freeevo, or full");
        }
    bool integration = false;
    if (p.boolean("integration", integration) && integration) {
        if (r.calculation != MSG_CALC_TIMEEVOLUTION)
            throw invalid("integration=true is a time-evolution output option; use calculation=yields for integrated "
                          "product/polarization observables");
        off("integration=true is not on the device path");
    }
    std::string ro = "haberkorn";
    p.str({"reactionoperators", "reactionoperator", "reaction_operator"}, ro);
    if (ro == "lindblad")
        throw invalid("reactionoperators=lindblad is a superspace reaction-superoperator model; HSGeneral supports "
                      "Haberkorn sink loss in Hilbert space");
    if (ro != "haberkorn") throw invalid("reactionoperators must be haberkorn in HSGeneral");
    std::string rm;
    if (p.str({"relaxationmodel", "relaxation_model"}, rm) && one_of(rm, {"nakajimazwanzig", "nakajima-zwanzig", "nz"}))
        throw invalid("Nakajima-Zwanzig relaxation is not available in HSGeneral; use the superspace NZ framework");

    std::string method;
    if (p.str({"method", "yieldmethod", "yield_method"}, method)) {
        if (r.calculation == MSG_CALC_YIELDS) {
            if (one_of(method, {"timeinf", "infinite", "timeinfinity", "time-infinity"})) out.time_infinity = 1;
            else if (!one_of(method, {"timeevo", "finite", "finitetime", "finite-time"}))
                throw invalid("yield method must be timeevo/finite or timeinf");
        } else if (method != "timeevo") {
            throw invalid("calculation=timeevolution requires method=timeevo when method is specified");
        }
    }
    if (dynamic && out.time_infinity)
        throw invalid("dynamic yields require finite-time propagation; method=timeinf is only defined for static "
                      "Hamiltonians");
    if (!offpath.empty()) throw unsupported(offpath);
}

}  // namespace

extern "C" msg_status msg_resolve_plan(const char *task_properties, msg_plan *plan, char *error, int32_t error_len) {
    auto put = [&](const std::string &m) {
        if (error && error_len > 0) {
            std::strncpy(error, m.c_str(), (size_t)error_len - 1);
            error[error_len - 1] = '\0';
        }
    };
    put("");
    if (!task_properties || !plan) {
        put("null argument");
        return MSG_INVALID_ARG;
    }
    try {
        resolve(Props(task_properties), *plan);
    } catch (const PlanError &e) {
        put(e.msg);
        return e.st;
    } catch (const std::exception &e) {
        put(e.what());
        return MSG_INVALID_ARG;
    }
    return MSG_OK;
}
