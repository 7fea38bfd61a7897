// MSD front-end: one MSD input file -> the structural descriptor of its spin
// system, the swept Zeeman field of every calculation step and a resolved plan for
// every Task, so that a whole RunSection sweep becomes one msg_propagate batch.
//
// Restates, for the objects of the Hilbert-space yield path:
//   - MSDParser::FileReader (MSDParser/FileReader.cpp:469-700): "//" and "/* */"
//     comments, quotes dropped with the spaces inside them kept, whitespace outside
//     strings dropped, `Keyword Name { contents }` objects inside `SpinSystem name`,
//     `Settings` and `Run` sections;
//   - ObjectParser property lists (msg_props.hpp);
//   - Spin (SpinAPI/Spin.cpp:29-60), Tensor::ParseTensor (SpinAPI/Tensor.cpp:219-300),
//     Interaction types (SpinAPI/Interaction.cpp:85-120), State kets
//     (SpinAPI/State.cpp:73-130, 225-330, 553-640), Transition (Transition.cpp:133-160,
//     234-242), SpinSystem initial states (SpinSystem.cpp:560-590);
//   - the step loop of main.cpp:600-680 with Action triggering (RunSection/Action.cpp:
//     100-140) and the RotateVector, FibonacciSphere and LogSpace actions
//     (RunSection/Actions/*.cpp), including their installation of point zero before
//     calculation step 1.
// Anything else (pulses, subsystems, trajectories, other actions or operators, more
// than one spin system, #include directives) is reported as MSG_UNSUPPORTED.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/molspin_gpu.h"
#include "msg_props.hpp"

using msg::lower;
using msg::Props;
using msg::trim;
using cd = std::complex<double>;

struct MsdError {
    msg_status st;
    std::string msg;
};

struct msg_msd {
    std::vector<msg_spin_desc> spins;
    std::vector<msg_interaction_desc> inter;
    std::vector<msg_state_desc> states;
    std::vector<msg_transition_desc> trans;
    std::vector<int32_t> initial;
    std::vector<msg_relaxation_desc> relax;
    std::vector<int32_t> pool;
    std::vector<double> amps;
    msg_system_desc desc{};
    std::string system_name;
    int32_t steps = 1;
    bool swept = false;          // an action drives the field of a Zeeman interaction
    std::vector<double> fields;  // steps x 3 (prefactor folded in) when swept
    struct Task {
        std::string name;
        msg_status st = MSG_OK;
        std::string why;
        msg_plan plan{};
    };
    std::vector<Task> tasks;
};

namespace {

// ---------------------------------------------------------------------------
// Reader
// ---------------------------------------------------------------------------
struct Object {
    std::string kind, name, body;
};
struct Section {
    std::string kind, name;
    std::vector<Object> objects;
};

class Reader {
  public:
    explicit Reader(const std::string &t) : s(t) {}

    std::vector<Section> sections() {
        std::vector<Section> out;
        for (;;) {
            skip();
            if (i >= s.size()) break;
            if (s[i] == '#') throw MsdError{MSG_UNSUPPORTED, "MSD directives (#include, #define) are not supported"};
            Section sec;
            sec.kind = lower(word());
            skip();
            if (i < s.size() && s[i] != '{') sec.name = word();
            expect('{');
            for (;;) {
                skip();
                if (i >= s.size()) throw MsdError{MSG_INVALID_ARG, "unterminated section " + sec.kind};
                if (s[i] == '}') {
                    ++i;
                    break;
                }
                Object o;
                o.kind = lower(word());
                skip();
                if (i < s.size() && s[i] != '{') o.name = word();
                expect('{');
                o.body = body();
                sec.objects.push_back(o);
            }
            out.push_back(sec);
        }
        return out;
    }

  private:
    const std::string &s;
    size_t i = 0;

    void skip() {  // whitespace and comments
        while (i < s.size()) {
            if (std::isspace((unsigned char)s[i])) ++i;
            else if (s.compare(i, 2, "//") == 0) {
                while (i < s.size() && s[i] != '\n') ++i;
            } else if (s.compare(i, 2, "/*") == 0) {
                const size_t e = s.find("*/", i + 2);
                i = e == std::string::npos ? s.size() : e + 2;
            } else break;
        }
    }
    std::string word() {
        std::string w;
        if (i < s.size() && s[i] == '"') {  // a quoted name
            const size_t e = s.find('"', i + 1);
            if (e == std::string::npos) throw MsdError{MSG_INVALID_ARG, "unterminated string"};
            w = s.substr(i + 1, e - i - 1);
            i = e + 1;
            return w;
        }
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == '.')) w += s[i++];
        if (w.empty()) throw MsdError{MSG_INVALID_ARG, std::string("unexpected character '") + (i < s.size() ? s[i] : ' ') +
                                                           "' in the MSD structure"};
        return w;
    }
    void expect(char c) {
        skip();
        if (i >= s.size() || s[i] != c) throw MsdError{MSG_INVALID_ARG, std::string("expected '") + c + "'"};
        ++i;
    }
    // Object contents up to the closing brace: quotes dropped (spaces inside kept),
    // whitespace outside strings dropped, comments skipped.
    std::string body() {
        std::string b;
        bool str = false;
        while (i < s.size()) {
            const char c = s[i];
            if (str) {
                if (c == '"') str = false;
                else b += c;
                ++i;
                continue;
            }
            if (c == '"') {
                str = true;
                ++i;
            } else if (s.compare(i, 2, "//") == 0 || s.compare(i, 2, "/*") == 0) {
                skip();
            } else if (c == '}') {
                ++i;
                return b;
            } else if (c == '{') {
                throw MsdError{MSG_INVALID_ARG, "nested braces inside an object"};
            } else {
                if (!std::isspace((unsigned char)c)) b += c;
                ++i;
            }
        }
        throw MsdError{MSG_INVALID_ARG, "unterminated object"};
    }
};

// ---------------------------------------------------------------------------
// Values
// ---------------------------------------------------------------------------
std::vector<double> numbers(const std::string &v) {  // space- or comma-separated
    std::vector<double> out;
    std::string t = v;
    for (char &c : t)
        if (c == ',') c = ' ';
    std::istringstream in(t);
    std::string tok;
    while (in >> tok) {
        try {
            out.push_back(std::stod(tok));
        } catch (...) {
            throw MsdError{MSG_INVALID_ARG, "not a number: " + tok};
        }
    }
    return out;
}

std::array<double, 3> vec3(const std::string &v, const std::string &what) {
    const std::vector<double> x = numbers(v);
    if (x.size() != 3) throw MsdError{MSG_INVALID_ARG, what + " needs three components"};
    return {x[0], x[1], x[2]};
}

std::vector<std::string> names(const std::string &v) {  // comma-separated list
    std::vector<std::string> out;
    std::istringstream in(v);
    for (std::string t; std::getline(in, t, ',');) {
        t = trim(t);
        if (!t.empty()) out.push_back(t);
    }
    return out;
}

// Spin quantum number as written ("1/2", "1", "3/2") -> multiplicity 2S+1.
int multiplicity(const std::string &v) {
    const std::string t = trim(v);
    int twice = 0;
    try {
        if (t.size() > 2 && t.substr(t.size() - 2) == "/2") twice = std::stoi(t.substr(0, t.size() - 2));
        else {
            const double x = 2.0 * std::stod(t);
            if (x != std::floor(x)) throw MsdError{MSG_INVALID_ARG, "spin must be a multiple of 1/2"};
            twice = (int)x;
        }
    } catch (const MsdError &) {
        throw;
    } catch (...) {
        throw MsdError{MSG_INVALID_ARG, "invalid spin quantum number " + t};
    }
    if (twice < 0) throw MsdError{MSG_INVALID_ARG, "negative spin quantum number"};
    return twice + 1;
}

// Tensor::ParseTensor: '+'-separated terms; a number or isotropic(x) adds to the
// isotropic value, anisotropic(a b c) to the principal values along axis1..3 (given
// by axis1(...) .. axis3(...), default x, y, z), matrix(...) sets a full matrix.
// Returns the lab-frame matrix (row-major) and whether it is isotropic.
bool tensor(const std::string &spec, double m[9]) {
    double iso = 0.0, an[3] = {0.0, 0.0, 0.0};
    double ax[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    bool has_matrix = false, has_aniso = false;
    double mat[9] = {0};
    std::istringstream in(spec);
    for (std::string term; std::getline(in, term, '+');) {
        term = trim(term);
        if (term.empty()) continue;
        const size_t j = term.find('(');
        if (j == std::string::npos) {
            iso += numbers(term).at(0);
            continue;
        }
        const size_t k = term.find(')', j);
        if (k == std::string::npos) throw MsdError{MSG_INVALID_ARG, "unbalanced parentheses in tensor " + spec};
        const std::string key = lower(term.substr(0, j)), val = term.substr(j + 1, k - j - 1);
        if (key == "isotropic") {
            iso += numbers(val).at(0);
        } else if (key == "anisotropic") {
            const std::vector<double> a = numbers(val);
            if (a.size() != 3) throw MsdError{MSG_INVALID_ARG, "anisotropic() needs three values"};
            for (int c = 0; c < 3; ++c) an[c] += a[c];
            has_aniso = true;
        } else if (key == "axis1" || key == "axis2" || key == "axis3") {
            const std::array<double, 3> a = vec3(val, key);
            const double n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            if (!(n > 0.0)) throw MsdError{MSG_INVALID_ARG, "zero tensor axis"};
            for (int c = 0; c < 3; ++c) ax[key[4] - '1'][c] = a[c] / n;
        } else if (key == "matrix") {
            std::vector<double> v;
            std::istringstream rows(val);
            for (std::string row; std::getline(rows, row, ';');) {
                const std::vector<double> r = numbers(row);
                v.insert(v.end(), r.begin(), r.end());
            }
            if (v.size() != 9) throw MsdError{MSG_INVALID_ARG, "matrix() needs 3 x 3 values"};
            for (int e = 0; e < 9; ++e) mat[e] = v[e];
            has_matrix = true;
        } else {
            throw MsdError{MSG_UNSUPPORTED, "tensor term " + key + " is not supported"};
        }
    }
    if (has_matrix && has_aniso) throw MsdError{MSG_UNSUPPORTED, "a tensor with both matrix() and anisotropic()"};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double v = (a == b ? iso : 0.0) + mat[3 * a + b];
            for (int p = 0; p < 3; ++p) v += an[p] * ax[p][a] * ax[p][b];
            m[3 * a + b] = v;
        }
    const double tr = (m[0] + m[4] + m[8]) / 3.0;
    double dev = 0.0;
    for (int e = 0; e < 9; ++e) dev += std::fabs(m[e] - (e % 4 == 0 ? tr : 0.0));
    return dev < 1e-10;
}

// A ket coefficient (State::ParseFactor): empty/+ -> 1, - -> -1, i / -i, a real
// number, or a number followed by i.
cd factor(const std::string &f) {
    const std::string t = trim(f);
    if (t.empty() || t == "+") return 1.0;
    if (t == "-") return -1.0;
    if (t == "i" || t == "+i" || t == "I" || t == "+I") return cd(0.0, 1.0);
    if (t == "-i" || t == "-I") return cd(0.0, -1.0);
    std::string num = t;
    bool imag = false;
    if (!num.empty() && (num.back() == 'i' || num.back() == 'I')) {
        imag = true;
        num.pop_back();
        if (num == "+" || num.empty()) num = "1";
        if (num == "-") num = "-1";
    }
    if (!num.empty() && num.back() == '*') num.pop_back();
    double x;
    try {
        size_t used = 0;
        x = std::stod(num, &used);
        if (used != num.size()) throw 0;
    } catch (...) {
        throw MsdError{MSG_UNSUPPORTED, "state coefficient " + t + " is not supported"};
    }
    return imag ? cd(0.0, x) : cd(x, 0.0);
}

// Amplitudes over the product basis of `mults` (first spin most significant) of a
// ket expression such as "|1/2,-1/2>-|-1/2,1/2>" or "0.5|1>+i|0>".
std::vector<cd> kets(const std::vector<int> &mults, const std::string &expr) {
    size_t n = 1;
    for (int m : mults) n *= (size_t)m;
    std::vector<cd> a(n, cd(0.0));
    size_t p = 0;
    while (p < expr.size()) {
        const size_t bar = expr.find('|', p);
        if (bar == std::string::npos) {
            if (!trim(expr.substr(p)).empty()) throw MsdError{MSG_INVALID_ARG, "text after the last ket: " + expr};
            break;
        }
        const cd c = factor(expr.substr(p, bar - p));
        const size_t close = expr.find('>', bar);
        if (close == std::string::npos) throw MsdError{MSG_INVALID_ARG, "unterminated ket in " + expr};
        const std::vector<std::string> ms = names(expr.substr(bar + 1, close - bar - 1));
        if (ms.size() != mults.size()) throw MsdError{MSG_INVALID_ARG, "ket " + expr.substr(bar, close - bar + 1) +
                                                                            " does not match the listed spins"};
        size_t idx = 0;
        for (size_t k = 0; k < ms.size(); ++k) {
            const int two_s = mults[k] - 1;
            int two_m;
            const std::string t = ms[k];
            if (t.size() > 2 && t.substr(t.size() - 2) == "/2") two_m = std::stoi(t.substr(0, t.size() - 2));
            else {
                const double x = 2.0 * std::stod(t);
                if (x != std::floor(x)) throw MsdError{MSG_INVALID_ARG, "invalid m in ket " + t};
                two_m = (int)x;
            }
            if (std::abs(two_m) > two_s || ((two_s - two_m) & 1)) throw MsdError{MSG_INVALID_ARG, "m out of range in ket"};
            idx = idx * (size_t)mults[k] + (size_t)((two_s - two_m) / 2);
        }
        a[idx] += c;
        p = close + 1;
    }
    return a;
}

// ---------------------------------------------------------------------------
// Actions (RunSection/Action.cpp and RunSection/Actions/*.cpp)
// ---------------------------------------------------------------------------
struct Target {
    int inter = -1;
    bool is_field = true;  // .field (vector) or .prefactor (scalar)
};

struct Act {
    std::string type;
    Target target;
    double value = 1.0;
    unsigned first = 1, last = 0, period = 1;
    // rotatevector
    double axis[3] = {0, 0, 1}, ref1[3] = {1, 0, 0}, ref2[3] = {0, 1, 0};
    // grids
    std::vector<std::array<double, 3>> vpoints;
    std::vector<double> spoints;
    size_t next = 0;
    bool triggered(unsigned step) const {  // Action::Step without looping
        return step >= first && !(last != 0 && step > last) && (step - first) % period == 0;
    }
};

// ActionFibonacciSphere point `i` of `n` (RunSection/Utility.cpp:44-79): y and the
// angle pass through float before the trigonometry, as in the reference.
std::array<double, 3> fib_point(int n, int i) {
    const double phi = M_PI * (3.0 - std::sqrt(5.0));
    const float y = (float)(1.0 - ((double)i / (double)(n - 1)) * 2.0);
    const float th = (float)(phi * i);
    const float yy = y * y;
    const double r = std::sqrt(1.0 - (double)yy);
    return {(double)std::cos(th) * r, (double)y, (double)std::sin(th) * r};
}

}  // namespace

// ---------------------------------------------------------------------------
// Compilation
// ---------------------------------------------------------------------------
static void compile(const std::string &text, msg_msd &m) {
    Reader rd(text);
    const std::vector<Section> secs = rd.sections();
    const Section *sys = nullptr;
    std::vector<const Section *> settings, runs;
    for (const Section &s : secs) {
        if (s.kind == "spinsystem") {
            if (sys) throw MsdError{MSG_UNSUPPORTED, "the device front-end takes one SpinSystem"};
            sys = &s;
        } else if (s.kind == "settings") settings.push_back(&s);
        else if (s.kind == "run") runs.push_back(&s);
        else throw MsdError{MSG_INVALID_ARG, "unknown MSD section " + s.kind};
    }
    if (!sys) throw MsdError{MSG_INVALID_ARG, "the MSD file defines no SpinSystem"};
    m.system_name = sys->name;
    // names are matched as written
    std::map<std::string, int> spin_of, state_of, inter_of;
    auto spin_idx = [&](const std::string &n) {
        auto it = spin_of.find(n);
        if (it == spin_of.end()) throw MsdError{MSG_INVALID_ARG, "unknown spin " + n};
        return it->second;
    };
    std::vector<int> mult;
    // spins first (objects may refer to spins declared later)
    for (const Object &o : sys->objects) {
        if (o.kind != "spin") continue;
        const Props p(o.body.c_str());
        msg_spin_desc d{};
        std::string v;
        const bool electron = p.str({"type"}, v) && (v == "electron" || v == "e");
        int mu = 2;
        if (p.text({"spin"}, v)) mu = multiplicity(v);
        d.multiplicity = mu;
        d.is_electron = electron ? 1 : 0;
        double t[9] = {2, 0, 0, 0, 2, 0, 0, 0, 2};  // default isotropic(2) (Spin.cpp:29)
        bool iso = true;
        if (p.text({"tensor"}, v)) iso = tensor(v, t);
        d.tensor_isotropic = iso ? 1 : 0;
        std::memcpy(d.tensor, t, sizeof t);
        spin_of[o.name] = (int)m.spins.size();
        m.spins.push_back(d);
        mult.push_back(mu);
    }
    // states
    for (const Object &o : sys->objects) {
        if (o.kind != "state") continue;
        const Props p(o.body.c_str());
        std::vector<int> sp;
        std::vector<cd> amp{cd(1.0)};
        for (const auto &kv : p.raw) {
            const std::string key = kv.first;
            const size_t j = key.find('(');
            const std::string fn = lower(key.substr(0, j == std::string::npos ? key.size() : j));
            if (j == std::string::npos || (fn != "spins" && fn != "spin"))
                throw MsdError{MSG_UNSUPPORTED, "State " + o.name + ": property " + key + " is not supported"};
            const size_t k = key.find(')', j);
            std::vector<int> here, mh;
            for (const std::string &n : names(key.substr(j + 1, k - j - 1))) {
                here.push_back(spin_idx(n));
                mh.push_back(mult[here.back()]);
            }
            const std::vector<cd> a = kets(mh, kv.second);
            std::vector<cd> prod(amp.size() * a.size());
            for (size_t x = 0; x < amp.size(); ++x)
                for (size_t y = 0; y < a.size(); ++y) prod[x * a.size() + y] = amp[x] * a[y];
            amp.swap(prod);
            sp.insert(sp.end(), here.begin(), here.end());
        }
        msg_state_desc d{};
        d.spins_offset = (int32_t)m.pool.size();
        d.spins_count = (int32_t)sp.size();
        d.amp_offset = (int32_t)(m.amps.size() / 2);
        m.pool.insert(m.pool.end(), sp.begin(), sp.end());
        for (const cd &c : amp) {
            m.amps.push_back(c.real());
            m.amps.push_back(c.imag());
        }
        state_of[o.name] = (int)m.states.size();
        m.states.push_back(d);
    }
    auto state_idx = [&](const std::string &n) {
        auto it = state_of.find(n);
        if (it == state_of.end()) throw MsdError{MSG_INVALID_ARG, "unknown state " + n};
        return it->second;
    };
    bool have_props = false;
    for (const Object &o : sys->objects) {
        const Props p(o.body.c_str());
        std::string v;
        if (o.kind == "spin" || o.kind == "state") continue;
        if (o.kind == "interaction") {
            msg_interaction_desc d{};
            p.str({"type"}, v);
            std::vector<std::string> g1, g2;
            std::string t;
            if (v == "onespin" || v == "singlespin" || v == "zeeman") {
                d.type = MSG_INTERACTION_ZEEMAN;
                if (!p.text({"spins", "group1"}, t)) throw MsdError{MSG_INVALID_ARG, "Interaction " + o.name + " lists no spins"};
                g1 = names(t);
                if (p.text({"field"}, t)) {
                    const std::array<double, 3> f = vec3(t, "field");
                    for (int a = 0; a < 3; ++a) d.field[a] = f[a];
                }
            } else if (v == "twospin" || v == "doublespin" || v == "hyperfine" || v == "exchange") {
                d.type = v == "exchange" ? MSG_INTERACTION_EXCHANGE : MSG_INTERACTION_DOUBLESPIN;
                if (!p.text({"group1"}, t)) throw MsdError{MSG_INVALID_ARG, "Interaction " + o.name + " has no group1"};
                g1 = names(t);
                if (!p.text({"group2"}, t)) throw MsdError{MSG_INVALID_ARG, "Interaction " + o.name + " has no group2"};
                g2 = names(t);
            } else {
                throw MsdError{MSG_UNSUPPORTED, "Interaction " + o.name + ": type " + v + " is not on the device path"};
            }
            if (p.text({"tensor"}, t)) {
                d.tensor_present = 1;
                d.tensor_isotropic = tensor(t, d.tensor) ? 1 : 0;
            }
            d.prefactor = 1.0;
            p.num("prefactor", d.prefactor);
            bool b = true;
            d.common_prefactor = p.boolean("commonprefactor", b) ? (b ? 1 : 0) : 1;
            b = false;
            d.ignore_tensors = p.boolean("ignoretensors", b) && b ? 1 : 0;
            d.td_frequency = 1.0;
            d.td_axis[2] = 1.0;
            if (p.str({"fieldtype", "timedependence"}, v)) {
                if (v == "linearpolarized" || v == "linearpolarization" || v == "oscillating" || v == "monochromatic")
                    d.time_dependence = MSG_TD_LINEAR;
                else if (v == "circularpolarized" || v == "circularpolarization")
                    d.time_dependence = MSG_TD_CIRCULAR;
                else if (v != "static" && v != "none")
                    throw MsdError{MSG_UNSUPPORTED, "Interaction " + o.name + ": field type " + v + " is not on the device path"};
                p.num("frequency", d.td_frequency);
                p.num("phase", d.td_phase);
                if (p.text({"axis"}, t)) {
                    const std::array<double, 3> a = vec3(t, "axis");
                    for (int c = 0; c < 3; ++c) d.td_axis[c] = a[c];
                }
                bool perp = false;
                d.td_perpendicular = p.boolean("perpendicularoscillations", perp) && perp ? 1 : 0;
            }
            d.group1_offset = (int32_t)m.pool.size();
            for (const std::string &n : g1) m.pool.push_back(spin_idx(n));
            d.group1_count = (int32_t)g1.size();
            d.group2_offset = (int32_t)m.pool.size();
            for (const std::string &n : g2) m.pool.push_back(spin_idx(n));
            d.group2_count = (int32_t)g2.size();
            inter_of[o.name] = (int)m.inter.size();
            m.inter.push_back(d);
        } else if (o.kind == "transition") {
            if (p.str({"type"}, v) && v != "sink")
                throw MsdError{MSG_UNSUPPORTED, "Transition " + o.name + ": only sinks are on the device path"};
            std::string src;
            if (!p.text({"sourcestate", "state", "source"}, src))
                throw MsdError{MSG_INVALID_ARG, "Transition " + o.name + " has no source state"};
            msg_transition_desc d{};
            d.source_state = state_idx(src);
            d.rate = 1.0;
            p.num("rate", d.rate);
            m.trans.push_back(d);
        } else if (o.kind == "operator") {
            if (!p.str({"type"}, v) || v != "relaxationdephasing")
                throw MsdError{MSG_UNSUPPORTED, "Operator " + o.name + ": only relaxationdephasing is on the device path"};
            std::string t;
            if (!p.text({"spins"}, t)) throw MsdError{MSG_INVALID_ARG, "Operator " + o.name + " lists no spins"};
            const std::vector<std::string> sp = names(t);
            if (sp.size() != 2) throw MsdError{MSG_UNSUPPORTED, "relaxationdephasing on two spins only"};
            msg_relaxation_desc d{};
            d.spin1 = spin_idx(sp[0]);
            d.spin2 = spin_idx(sp[1]);
            p.num("rate", d.rate);
            m.relax.push_back(d);
        } else if (o.kind == "properties") {
            if (have_props) throw MsdError{MSG_INVALID_ARG, "a second Properties object"};
            have_props = true;
            std::string t;
            if (p.text({"initialstate"}, t))
                for (const std::string &n : names(t)) {
                    if (lower(n) == "thermal") throw MsdError{MSG_UNSUPPORTED, "thermal initial states"};
                    m.initial.push_back(state_idx(n));
                }
        } else {
            throw MsdError{MSG_UNSUPPORTED, "SpinSystem object type " + o.kind + " is not on the device path"};
        }
    }
    // Settings: steps and actions
    std::vector<Act> acts;
    auto target = [&](const std::string &name, bool vector) {
        const size_t a = name.find('.'), b = name.rfind('.');
        if (a == std::string::npos || a == b) throw MsdError{MSG_INVALID_ARG, "action target " + name + " is not system.object.field"};
        if (name.substr(0, a) != m.system_name) throw MsdError{MSG_INVALID_ARG, "action target " + name + " names another system"};
        auto it = inter_of.find(name.substr(a + 1, b - a - 1));
        const std::string what = lower(name.substr(b + 1));
        if (it == inter_of.end() || (vector && what != "field") || (!vector && what != "prefactor"))
            throw MsdError{MSG_UNSUPPORTED, "action target " + name + " is not a Zeeman field or prefactor"};
        if (m.inter[it->second].type != MSG_INTERACTION_ZEEMAN)
            throw MsdError{MSG_UNSUPPORTED, "action target " + name + ": only Zeeman interactions can be swept"};
        if (m.inter[it->second].time_dependence != MSG_TD_NONE)
            throw MsdError{MSG_UNSUPPORTED, "action target " + name + " has a time-dependent field"};
        return Target{it->second, vector};
    };
    // field and prefactor of every interaction while the actions run
    std::vector<std::array<double, 3>> field(m.inter.size());
    std::vector<double> pref(m.inter.size());
    for (size_t k = 0; k < m.inter.size(); ++k) {
        for (int a = 0; a < 3; ++a) field[k][a] = m.inter[k].field[a];
        pref[k] = m.inter[k].prefactor;
    }
    for (const Section *s : settings)
        for (const Object &o : s->objects) {
            const Props p(o.body.c_str());
            if (o.kind == "settings") {
                int st = 1;
                if (p.integer("steps", st)) {
                    if (st < 1) throw MsdError{MSG_INVALID_ARG, "steps must be positive"};
                    m.steps = st;
                }
            } else if (o.kind == "action") {
                Act a;
                if (!p.str({"type"}, a.type)) throw MsdError{MSG_INVALID_ARG, "Action " + o.name + " has no type"};
                p.num("value", a.value);
                int x;
                if (p.integer("first", x)) a.first = (unsigned)x;
                if (p.integer("last", x)) a.last = (unsigned)x;
                if (p.integer("period", x)) a.period = (unsigned)x;
                bool loop = false;
                if (p.boolean("loop", loop) && loop) throw MsdError{MSG_UNSUPPORTED, "looping actions"};
                if (a.period == 0) throw MsdError{MSG_INVALID_ARG, "Action " + o.name + " has period=0"};
                std::string t;
                if (a.type == "rotatevector") {
                    if (!p.text({"actionvector", "vector"}, t)) throw MsdError{MSG_INVALID_ARG, "RotateVector without a vector"};
                    a.target = target(t, true);
                    std::string axs;
                    if (!p.text({"axis"}, axs)) throw MsdError{MSG_INVALID_ARG, "RotateVector without an axis"};
                    const std::array<double, 3> v = vec3(axs, "axis");
                    const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    if (!(n > 0.0)) throw MsdError{MSG_INVALID_ARG, "zero rotation axis"};
                    for (int c = 0; c < 3; ++c) a.axis[c] = v[c] / n;
                    // reference axis: the coordinate axis least aligned with the rotation axis,
                    // made orthogonal to it; the second one completes a right-handed frame
                    int mi = 0;
                    for (int c = 1; c < 3; ++c)
                        if (std::fabs(a.axis[c]) < std::fabs(a.axis[mi])) mi = c;
                    double r1[3] = {0, 0, 0};
                    r1[mi] = 1.0;
                    for (int c = 0; c < 3; ++c) r1[c] -= a.axis[mi] * a.axis[c];
                    const double n1 = std::sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
                    for (int c = 0; c < 3; ++c) a.ref1[c] = r1[c] / n1;
                    double r2[3] = {a.axis[1] * a.ref1[2] - a.axis[2] * a.ref1[1], a.axis[2] * a.ref1[0] - a.axis[0] * a.ref1[2],
                                    a.axis[0] * a.ref1[1] - a.axis[1] * a.ref1[0]};
                    const double n2 = std::sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
                    for (int c = 0; c < 3; ++c) a.ref2[c] = r2[c] / n2;
                } else if (a.type == "fibonaccisphere") {
                    if (!p.text({"actionvector", "vector"}, t)) throw MsdError{MSG_INVALID_ARG, "FibonacciSphere without a vector"};
                    a.target = target(t, true);
                    int n = 0;
                    if (!p.integer("points", n) || n < 2) throw MsdError{MSG_INVALID_ARG, "FibonacciSphere needs points >= 2"};
                    const std::array<double, 3> &v = field[a.target.inter];
                    const double mag = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    for (int i = 0; i < n; ++i) {
                        std::array<double, 3> q = fib_point(n, i);
                        for (double &c : q) c *= mag;
                        a.vpoints.push_back(q);
                    }
                } else if (a.type == "logspace") {
                    if (!p.text({"actionscalar", "scalar"}, t)) throw MsdError{MSG_INVALID_ARG, "LogSpace without a scalar"};
                    a.target = target(t, false);
                    int n = 0;
                    double lo, hi;
                    if (!p.integer("points", n) || n < 1) throw MsdError{MSG_INVALID_ARG, "LogSpace needs points >= 1"};
                    if (!p.num("minvalue", lo) && !p.num("min", lo)) throw MsdError{MSG_INVALID_ARG, "LogSpace without a minimum"};
                    if (!p.num("maxvalue", hi) && !p.num("max", hi)) throw MsdError{MSG_INVALID_ARG, "LogSpace without a maximum"};
                    if (!(lo < hi)) throw MsdError{MSG_INVALID_ARG, "LogSpace bounds"};
                    for (int i = 0; i < n; ++i)  // arma::logspace
                        a.spoints.push_back(n == 1 ? std::pow(10.0, hi) : std::pow(10.0, lo + (hi - lo) * i / (n - 1)));
                } else {
                    throw MsdError{MSG_UNSUPPORTED, "Action type " + a.type + " is not supported by the front-end"};
                }
                // grid actions install point zero before step 1 when they start there
                if (a.first == 1 && (!a.vpoints.empty() || !a.spoints.empty())) {
                    if (!a.vpoints.empty()) field[a.target.inter] = a.vpoints[a.next++];
                    else pref[a.target.inter] = a.spoints[a.next++];
                }
                acts.push_back(a);
            } else if (o.kind != "output") {
                throw MsdError{MSG_UNSUPPORTED, "Settings object type " + o.kind + " is not supported"};
            }
        }
    if (!acts.empty()) {
        const int k = acts[0].target.inter;
        for (const Act &a : acts)
            if (a.target.inter != k) throw MsdError{MSG_UNSUPPORTED, "actions on more than one interaction"};
        m.swept = true;
        m.inter[k].field_swept = 1;
        // the swept field carries the prefactor: field_item = prefactor x field
        m.fields.assign((size_t)3 * m.steps, 0.0);
        for (int step = 1; step <= m.steps; ++step) {
            for (int c = 0; c < 3; ++c) m.fields[3 * (size_t)(step - 1) + c] = pref[k] * field[k][c];
            for (Act &a : acts) {  // actions after each calculation step (main.cpp:678)
                if (!a.triggered((unsigned)step + 1)) continue;
                if (a.type == "rotatevector") {
                    std::array<double, 3> &v = field[k];
                    const double pr = a.axis[0] * v[0] + a.axis[1] * v[1] + a.axis[2] * v[2];
                    double w[3];
                    for (int c = 0; c < 3; ++c) w[c] = v[c] - pr * a.axis[c];
                    const double len = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                    const double p1 = a.ref1[0] * w[0] + a.ref1[1] * w[1] + a.ref1[2] * w[2];
                    const double p2 = a.ref2[0] * w[0] + a.ref2[1] * w[1] + a.ref2[2] * w[2];
                    const double ang = std::atan2(p2, p1) + a.value / 180.0 * M_PI;
                    for (int c = 0; c < 3; ++c)
                        v[c] = len * (a.ref1[c] * std::cos(ang) + a.ref2[c] * std::sin(ang)) + a.axis[c] * pr;
                } else if (!a.vpoints.empty()) {
                    if (a.next < a.vpoints.size()) field[k] = a.vpoints[a.next++];
                } else if (a.next < a.spoints.size()) {
                    pref[k] = a.spoints[a.next++];
                }
            }
        }
        for (int c = 0; c < 3; ++c) m.inter[k].field[c] = m.fields[c];
        m.inter[k].prefactor = 1.0;
    }
    // Run: every task through the plan resolver
    for (const Section *s : runs)
        for (const Object &o : s->objects) {
            if (o.kind != "task") throw MsdError{MSG_UNSUPPORTED, "Run object type " + o.kind + " is not supported"};
            msg_msd::Task t;
            t.name = o.name;
            char why[512];
            t.st = msg_resolve_plan(o.body.c_str(), &t.plan, why, sizeof why);
            t.why = why;
            m.tasks.push_back(t);
        }
    // descriptor
    msg_system_desc &d = m.desc;
    d = msg_system_desc{};
    d.n_spins = (int32_t)m.spins.size();
    d.spins = m.spins.data();
    d.n_interactions = (int32_t)m.inter.size();
    d.interactions = m.inter.data();
    d.n_states = (int32_t)m.states.size();
    d.states = m.states.data();
    d.n_transitions = (int32_t)m.trans.size();
    d.transitions = m.trans.data();
    d.n_initial = (int32_t)m.initial.size();
    d.initial_states = m.initial.data();
    d.initial_weights = nullptr;
    d.n_relaxations = (int32_t)m.relax.size();
    d.relaxations = m.relax.data();
    d.index_pool_len = (int32_t)m.pool.size();
    d.index_pool = m.pool.data();
    d.amp_pool_len = (int32_t)(m.amps.size() / 2);
    d.amp_pool = m.amps.data();
}

static void put(char *dst, int32_t len, const std::string &s) {
    if (!dst || len <= 0) return;
    std::strncpy(dst, s.c_str(), (size_t)len - 1);
    dst[len - 1] = '\0';
}

extern "C" {

msg_status msg_msd_compile(const char *text, msg_msd **out, char *error, int32_t error_len) {
    put(error, error_len, "");
    if (!text || !out) {
        put(error, error_len, "null argument");
        return MSG_INVALID_ARG;
    }
    *out = nullptr;
    std::unique_ptr<msg_msd> m(new msg_msd());
    try {
        compile(text, *m);
    } catch (const MsdError &e) {
        put(error, error_len, e.msg);
        return e.st;
    } catch (const std::exception &e) {
        put(error, error_len, std::string("invalid MSD input: ") + e.what());
        return MSG_INVALID_ARG;
    }
    *out = m.release();
    return MSG_OK;
}

msg_status msg_msd_system(const msg_msd *m, const msg_system_desc **desc) {
    if (!m || !desc) return MSG_INVALID_ARG;
    *desc = &m->desc;
    return MSG_OK;
}

msg_status msg_msd_steps(const msg_msd *m, int32_t *n_steps, int32_t *swept, double *fields) {
    if (!m || !n_steps) return MSG_INVALID_ARG;
    *n_steps = m->steps;
    if (swept) *swept = m->swept ? 1 : 0;
    if (fields && m->swept) std::copy(m->fields.begin(), m->fields.end(), fields);
    return MSG_OK;
}

msg_status msg_msd_task_count(const msg_msd *m, int32_t *n) {
    if (!m || !n) return MSG_INVALID_ARG;
    *n = (int32_t)m->tasks.size();
    return MSG_OK;
}

msg_status msg_msd_task(const msg_msd *m, int32_t i, msg_plan *plan, char *name, int32_t name_len, char *why,
                        int32_t why_len) {
    if (!m || i < 0 || i >= (int32_t)m->tasks.size()) return MSG_INVALID_ARG;
    const msg_msd::Task &t = m->tasks[(size_t)i];
    if (plan) *plan = t.plan;
    put(name, name_len, t.name);
    put(why, why_len, t.why);
    return t.st;
}

msg_status msg_msd_destroy(msg_msd *m) {
    delete m;
    return MSG_OK;
}

}  // extern "C"
