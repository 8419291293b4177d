// paras.cpp — runtime reader of the reference's parameter headers.
//
// The reference configures every run at compile time: paras/paraK.h and
// paras/total.h hold lines such as
//     const highprecision dx=0.03,dy=0.03,dxdy=dx*dy;
//     const int dimX=512,dimY=512;
//     float R1=50,R2=R1/2;
//     int timesteps=500;
//     #define Monitor2
// (snowflake/paras/para1.h:4-15, total.h:18-47; graingrowth/paras/para1.h:4-6;
// sintering/paras/para1.h:1-9) and its Python harness rewrites those lines
// and recompiles (snowflake/end2end_mix.py:23-61).  This reader evaluates the
// same declarations at run time (C arithmetic: integer division between
// integers) and maps the names each application uses onto amps_config.
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "amps_internal.hpp"

namespace {

thread_local std::string g_parse_error;

struct Val {
    double v = 0;
    bool is_int = false;
};

struct Parser {
    std::string s;
    size_t i = 0;
    const std::map<std::string, Val>& sym;
    bool ok = true;
    Parser(const std::string& str, const std::map<std::string, Val>& m) : s(str), sym(m) {}

    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    Val expr() {
        Val a = term();
        for (;;) {
            ws();
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
                char op = s[i++];
                Val b = term();
                a = arith(a, b, op);
            } else return a;
        }
    }
    // An integer division by zero is a parse error (C leaves it undefined).
    Val divide_checked(Val a, Val b) {
        if (a.is_int && b.is_int && b.v == 0) {
            ok = false;
            g_parse_error = "integer division by zero";
            return {};
        }
        return arith(a, b, '/');
    }
    Val term() {
        Val a = factor();
        for (;;) {
            ws();
            if (i < s.size() && (s[i] == '*' || s[i] == '/')) {
                char op = s[i++];
                Val b = factor();
                a = op == '/' ? divide_checked(a, b) : arith(a, b, op);
            } else return a;
        }
    }
    Val factor() {
        ws();
        if (i >= s.size()) { ok = false; return {}; }
        if (s[i] == '(') {
            ++i;
            Val v = expr();
            ws();
            if (i < s.size() && s[i] == ')') ++i; else ok = false;
            return v;
        }
        if (s[i] == '-' || s[i] == '+') {
            char op = s[i++];
            Val v = factor();
            if (op == '-') v.v = -v.v;
            return v;
        }
        if (std::isdigit((unsigned char)s[i]) || s[i] == '.') {
            size_t j = i;
            bool flt = false;
            while (j < s.size() && (std::isdigit((unsigned char)s[j]) || s[j] == '.' || s[j] == 'e' || s[j] == 'E' ||
                                    ((s[j] == '-' || s[j] == '+') && j > i && (s[j - 1] == 'e' || s[j - 1] == 'E')))) {
                if (s[j] == '.' || s[j] == 'e' || s[j] == 'E') flt = true;
                ++j;
            }
            Val v;
            v.v = std::strtod(s.substr(i, j - i).c_str(), nullptr);
            v.is_int = !flt;
            i = j;
            while (i < s.size() && (s[i] == 'f' || s[i] == 'F' || s[i] == 'l' || s[i] == 'L')) { v.is_int = false; ++i; }
            return v;
        }
        if (std::isalpha((unsigned char)s[i]) || s[i] == '_') {
            size_t j = i;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
            std::string name = s.substr(i, j - i);
            i = j;
            auto it = sym.find(name);
            if (it == sym.end()) { ok = false; g_parse_error = "unknown name '" + name + "'"; return {}; }
            return it->second;
        }
        ok = false;
        return {};
    }
    static Val arith(Val a, Val b, char op) {
        Val r;
        r.is_int = a.is_int && b.is_int;
        switch (op) {
            case '+': r.v = a.v + b.v; break;
            case '-': r.v = a.v - b.v; break;
            case '*': r.v = a.v * b.v; break;
            default:
                if (r.is_int) r.v = (double)((long long)a.v / (long long)b.v);  // b != 0: divide_checked
                else r.v = a.v / b.v;
        }
        return r;
    }
};

std::string strip_comments(const std::string& in) {
    std::string out;
    for (size_t i = 0; i < in.size();) {
        if (in.compare(i, 2, "//") == 0) {
            while (i < in.size() && in[i] != '\n') ++i;
        } else if (in.compare(i, 2, "/*") == 0) {
            size_t e = in.find("*/", i + 2);
            i = e == std::string::npos ? in.size() : e + 2;
        } else {
            out += in[i++];
        }
    }
    return out;
}

bool is_type_word(const std::string& w) {
    return w == "const" || w == "int" || w == "float" || w == "double" || w == "highprecision" ||
           w == "long" || w == "unsigned" || w == "static";
}

void apply(amps_config* cfg, const std::map<std::string, Val>& sym, int* timesteps, bool* set_ts) {
    auto get = [&](const char* n, double* dst) {
        auto it = sym.find(n);
        if (it != sym.end()) { *dst = it->second.v; return true; }
        return false;
    };
    double v;
    if (get("dimX", &v)) cfg->nx = (int)v;
    if (get("dimY", &v)) cfg->ny = (int)v;
    if (get("nx", &v)) cfg->nx = (int)v;
    if (get("ny", &v)) cfg->ny = (int)v;
    if (get("unitx", &v)) cfg->tile = (int)v;
    if (get("tile", &v)) cfg->tile = (int)v;
    if (get("threshold", &v)) cfg->threshold = v;
    if (get("monitor_every", &v)) cfg->monitor_every = (int)v;
    if (get("monitor_field", &v)) cfg->monitor_field = (int)v;
    if (get("Q", &v) || get("q", &v)) cfg->q = (int)v;
    if (get("rng_seed", &v)) cfg->seed = (uint64_t)v;
    if (get("timesteps", &v)) { if (timesteps) *timesteps = (int)v; *set_ts = true; }
    double* p = cfg->params;
    if (cfg->model == AMPS_SNOWFLAKE) {
        // snowflake/paras/para1.h:4-15; nucleus centre (dimX/2+8, dimY/2+8) from tools.h:27
        get("dx", &p[AMPS_SF_DX]); get("dy", &p[AMPS_SF_DY]); get("dtime", &p[AMPS_SF_DT]);
        get("tau", &p[AMPS_SF_TAU]); get("epsilonb", &p[AMPS_SF_EPSB]); get("kappa", &p[AMPS_SF_KAPPA]);
        get("delta", &p[AMPS_SF_DELTA]); get("aniso", &p[AMPS_SF_ANISO]); get("alpha", &p[AMPS_SF_ALPHA]);
        get("gama", &p[AMPS_SF_GAMMA]); get("teq", &p[AMPS_SF_TEQ]); get("theta0", &p[AMPS_SF_THETA0]);
        get("pi", &p[AMPS_SF_PI]); get("seed", &p[AMPS_SF_R2]); get("noise", &p[AMPS_SF_NOISE]);
        if (!get("cx", &p[AMPS_SF_CX])) p[AMPS_SF_CX] = cfg->nx / 2 + 8;
        if (!get("cy", &p[AMPS_SF_CY])) p[AMPS_SF_CY] = cfg->ny / 2 + 8;
    } else if (cfg->model == AMPS_GRAINGROWTH) {
        // graingrowth/paras/para1.h:4-6
        get("dtime", &p[AMPS_GG_DT]); get("mobil", &p[AMPS_GG_L]); get("dx", &p[AMPS_GG_DX]);
        get("dy", &p[AMPS_GG_DY]); get("grcoef", &p[AMPS_GG_KAPPA]); get("R", &p[AMPS_GG_R]);
        if (!get("Rx", &p[AMPS_GG_CX])) p[AMPS_GG_CX] = cfg->nx / 2;
        if (!get("Ry", &p[AMPS_GG_CY])) p[AMPS_GG_CY] = cfg->ny / 2;
        get("alpha", &p[AMPS_GG_ALPHA]); get("beta", &p[AMPS_GG_BETA]); get("gamma", &p[AMPS_GG_GAMMA]);
        get("amp", &p[AMPS_GG_AMP]);
    } else {
        // sintering/paras/para1.h:5-7; A=16, B=1 are literals in sintering/function.h:72
        get("coefm", &p[AMPS_SN_KC]); get("coefk", &p[AMPS_SN_KE]); get("coefl", &p[AMPS_SN_L]);
        get("dvol", &p[AMPS_SN_DVOL]); get("dvap", &p[AMPS_SN_DVAP]); get("dsur", &p[AMPS_SN_DSURF]);
        get("dgrb", &p[AMPS_SN_DGB]); get("dx", &p[AMPS_SN_DX]); get("dy", &p[AMPS_SN_DY]);
        get("dtime", &p[AMPS_SN_DT]); get("R1", &p[AMPS_SN_R1]); get("R2", &p[AMPS_SN_R2]);
        get("A", &p[AMPS_SN_A]); get("B", &p[AMPS_SN_B]);
        if (get("Rx1", &p[AMPS_SN_X1])) p[AMPS_SN_X2] = p[AMPS_SN_X1];
        get("Rx2", &p[AMPS_SN_X2]);
        get("Ry1", &p[AMPS_SN_Y1]); get("Ry2", &p[AMPS_SN_Y2]);
        get("width", &p[AMPS_SN_WIDTH]); get("noise", &p[AMPS_SN_NOISE]);
    }
}

}  // namespace

extern "C" int amps_load_paras(const char* path, amps_config* cfg, int* timesteps) {
    if (!path || !cfg) return AMPS_E_ARG;
    std::ifstream f(path);
    if (!f) { amps::set_error_text((std::string("cannot open ") + path).c_str()); return AMPS_E_IO; }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = strip_comments(ss.str());
    std::map<std::string, Val> sym;
    // names provided by earlier headers of the reference (total.h)
    sym["dimX"] = {(double)cfg->nx, true};
    sym["dimY"] = {(double)cfg->ny, true};
    std::istringstream lines(text);
    std::string line;
    int lineno = 0;
    // statements may span lines: accumulate up to ';'
    std::string stmt;
    while (std::getline(lines, line)) {
        ++lineno;
        size_t a = line.find_first_not_of(" \t\r");
        if (a == std::string::npos) continue;
        std::string t = line.substr(a);
        if (t[0] == '#') {
            std::istringstream ws(t.substr(1));
            std::string kw, name, val;
            ws >> kw >> name;
            if (kw == "define") {
                std::getline(ws, val);
                size_t b = val.find_first_not_of(" \t");
                if (b != std::string::npos) {
                    Parser P(val.substr(b), sym);
                    Val v = P.expr();
                    if (P.ok) sym[name] = v;
                }
                if (name == "AMSTENCIL") cfg->mode = AMPS_ADAPTIVE;
                else if (name == "PURE") cfg->mode = AMPS_PURE_FP64;
                else if (name == "GRAM1" || name == "GRAM2") cfg->mode = AMPS_STATIC_MAP;
                else if (name == "Monitor1") cfg->monitor = AMPS_MON_SPATIAL;
                else if (name == "Monitor2") cfg->monitor = AMPS_MON_TEMPORAL;
                else if (name == "MonitorGradient") cfg->monitor = AMPS_MON_GRADIENT;
                else if (name == "NEUMANN") cfg->bc = AMPS_BC_NEUMANN;
                else if (name == "LAP5") cfg->stencil = AMPS_STENCIL_LAP5;
                else if (name == "INIT_CONFIG") cfg->init = AMPS_INIT_CONFIG;
            }
            continue;
        }
        stmt += " " + t;
        size_t semi;
        while ((semi = stmt.find(';')) != std::string::npos) {
            std::string st = stmt.substr(0, semi);
            stmt = stmt.substr(semi + 1);
            // tokens before the first '=': type words then a name
            std::istringstream ws(st);
            std::string w;
            size_t pos = st.find_first_not_of(" \t");
            if (pos == std::string::npos) continue;
            if (st.compare(pos, 7, "typedef") == 0 || st.compare(pos, 5, "using") == 0) continue;
            // skip type words
            bool is_int = false;
            for (;;) {
                size_t e = pos;
                while (e < st.size() && (std::isalnum((unsigned char)st[e]) || st[e] == '_')) ++e;
                std::string word = st.substr(pos, e - pos);
                if (!is_type_word(word)) break;
                if (word == "int" || word == "long" || word == "unsigned") is_int = true;
                pos = st.find_first_not_of(" \t", e);
                if (pos == std::string::npos) break;
            }
            if (pos == std::string::npos) continue;
            // declarators separated by top-level commas
            std::string rest = st.substr(pos);
            int depth = 0;
            size_t s0 = 0;
            for (size_t k = 0; k <= rest.size(); ++k) {
                if (k < rest.size() && rest[k] == '(') depth++;
                if (k < rest.size() && rest[k] == ')') depth--;
                if (k == rest.size() || (rest[k] == ',' && depth == 0)) {
                    std::string decl = rest.substr(s0, k - s0);
                    s0 = k + 1;
                    size_t eq = decl.find('=');
                    if (eq == std::string::npos) continue;
                    std::string name = decl.substr(0, eq);
                    name.erase(0, name.find_first_not_of(" \t"));
                    name.erase(name.find_last_not_of(" \t") + 1);
                    Parser P(decl.substr(eq + 1), sym);
                    Val v = P.expr();
                    P.ws();
                    if (!P.ok || P.i != P.s.size()) {
                        std::string msg = "line " + std::to_string(lineno) + ": cannot evaluate '" + decl + "'";
                        if (!g_parse_error.empty()) msg += " (" + g_parse_error + ")";
                        g_parse_error.clear();
                        amps::set_error_text(msg.c_str());
                        return AMPS_E_PARSE;
                    }
                    if (is_int) { v.v = (double)(long long)v.v; v.is_int = true; }
                    else v.is_int = false;
                    sym[name] = v;
                }
            }
        }
    }
    bool set_ts = false;
    apply(cfg, sym, timesteps, &set_ts);
    return AMPS_OK;
}
