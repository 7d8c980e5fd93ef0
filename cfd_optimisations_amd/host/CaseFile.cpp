// CaseFile.cpp -- INI case files for hcfd_run (see CaseFile.hpp).
#include "CaseFile.hpp"

#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "hcfd.h"

namespace casefile {

// ---- expressions -------------------------------------------------------------

struct Expr::Node {
    enum Kind { NUM, X, Y, NEG, BIN, CALL } kind = NUM;
    double value = 0.0;
    char op = 0;               // + - * / ^ < > l(<=) g(>=) e(==) n(!=)
    std::string fn;
    std::vector<std::shared_ptr<const Node>> args;
};

namespace {

using NodeP = std::shared_ptr<const Expr::Node>;

class Parser {
   public:
    explicit Parser(const std::string& s) : s(s) {}
    NodeP parse() {
        NodeP n = comparison();
        skip();
        if (i != s.size()) fail("unexpected '" + s.substr(i, 1) + "'");
        return n;
    }

   private:
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string& why) const {
        throw std::runtime_error("expression '" + s + "': " + why + " at position " + std::to_string(i));
    }
    void skip() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool eat(const char* t) {
        skip();
        const size_t n = std::char_traits<char>::length(t);
        if (s.compare(i, n, t) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    static NodeP bin(char op, NodeP a, NodeP b) {
        auto n = std::make_shared<Expr::Node>();
        n->kind = Expr::Node::BIN;
        n->op = op;
        n->args = {a, b};
        return n;
    }
    NodeP comparison() {
        NodeP a = sum();
        for (;;) {
            if (eat("<=")) a = bin('l', a, sum());
            else if (eat(">=")) a = bin('g', a, sum());
            else if (eat("==")) a = bin('e', a, sum());
            else if (eat("!=")) a = bin('n', a, sum());
            else if (eat("<")) a = bin('<', a, sum());
            else if (eat(">")) a = bin('>', a, sum());
            else return a;
        }
    }
    NodeP sum() {
        NodeP a = product();
        for (;;) {
            if (eat("+")) a = bin('+', a, product());
            else if (eat("-")) a = bin('-', a, product());
            else return a;
        }
    }
    NodeP product() {
        NodeP a = unary();
        for (;;) {
            if (eat("*")) a = bin('*', a, unary());
            else if (eat("/")) a = bin('/', a, unary());
            else return a;
        }
    }
    NodeP unary() {
        if (eat("-")) {
            auto n = std::make_shared<Expr::Node>();
            n->kind = Expr::Node::NEG;
            n->args = {unary()};
            return n;
        }
        if (eat("+")) return unary();
        return power();
    }
    NodeP power() {
        NodeP a = atom();
        if (eat("^")) return bin('^', a, unary());  // right-associative, binds tighter than unary minus on the left
        return a;
    }
    NodeP atom() {
        skip();
        if (i >= s.size()) fail("unexpected end");
        if (eat("(")) {
            NodeP a = comparison();
            if (!eat(")")) fail("')' expected");
            return a;
        }
        const char c = s[i];
        if (std::isdigit((unsigned char)c) || c == '.') {
            size_t used = 0;
            double v = 0.0;
            try {
                v = std::stod(s.substr(i), &used);
            } catch (const std::exception&) {
                fail("bad number");
            }
            i += used;
            auto n = std::make_shared<Expr::Node>();
            n->value = v;
            return n;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) j++;
            const std::string id = s.substr(i, j - i);
            i = j;
            auto n = std::make_shared<Expr::Node>();
            if (id == "x") n->kind = Expr::Node::X;
            else if (id == "y") n->kind = Expr::Node::Y;
            else if (id == "pi") n->value = M_PI;
            else {
                static const char* fns[] = {"sqrt", "exp", "log", "sin", "cos", "tan", "abs", "min", "max", "pow"};
                bool known = false;
                for (const char* f : fns) known = known || id == f;
                if (!known) fail("unknown name '" + id + "'");
                if (!eat("(")) fail("'(' expected after " + id);
                n->kind = Expr::Node::CALL;
                n->fn = id;
                n->args.push_back(comparison());
                while (eat(",")) n->args.push_back(comparison());
                if (!eat(")")) fail("')' expected");
                const size_t want = (id == "min" || id == "max" || id == "pow") ? 2 : 1;
                if (n->args.size() != want) fail(id + " takes " + std::to_string(want) + " argument(s)");
            }
            return n;
        }
        fail("unexpected '" + std::string(1, c) + "'");
    }
};

double eval(const Expr::Node& n, double x, double y) {
    switch (n.kind) {
        case Expr::Node::NUM: return n.value;
        case Expr::Node::X: return x;
        case Expr::Node::Y: return y;
        case Expr::Node::NEG: return -eval(*n.args[0], x, y);
        case Expr::Node::BIN: {
            const double a = eval(*n.args[0], x, y), b = eval(*n.args[1], x, y);
            switch (n.op) {
                case '+': return a + b;
                case '-': return a - b;
                case '*': return a * b;
                case '/': return a / b;
                case '^': return std::pow(a, b);
                case '<': return a < b ? 1.0 : 0.0;
                case '>': return a > b ? 1.0 : 0.0;
                case 'l': return a <= b ? 1.0 : 0.0;
                case 'g': return a >= b ? 1.0 : 0.0;
                case 'e': return a == b ? 1.0 : 0.0;
                default: return a != b ? 1.0 : 0.0;
            }
        }
        case Expr::Node::CALL: {
            const double a = eval(*n.args[0], x, y);
            const std::string& f = n.fn;
            if (f == "sqrt") return std::sqrt(a);
            if (f == "exp") return std::exp(a);
            if (f == "log") return std::log(a);
            if (f == "sin") return std::sin(a);
            if (f == "cos") return std::cos(a);
            if (f == "tan") return std::tan(a);
            if (f == "abs") return std::fabs(a);
            const double b = eval(*n.args[1], x, y);
            if (f == "min") return std::fmin(a, b);
            if (f == "max") return std::fmax(a, b);
            return std::pow(a, b);
        }
    }
    return 0.0;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

}  // namespace

Expr::Expr(const std::string& text) : src(text), root(Parser(text).parse()) {}

double Expr::operator()(double x, double y) const { return eval(*root, x, y); }

int bc_type_of(const std::string& name) {
    const std::string n = lower(name);
    if (n == "symmetry") return HCFD_SYMMETRY;
    if (n == "wall") return HCFD_WALL;
    if (n == "sup-inlet") return HCFD_SUPINLET;
    if (n == "sub-inlet") return HCFD_SUBINLET;
    if (n == "sup-outlet") return HCFD_SUPOUTLET;
    if (n == "sub-outlet") return HCFD_SUBOUTLET;
    throw std::runtime_error("unknown BC type '" + name +
                             "' (symmetry, wall, sup-inlet, sub-inlet, sup-outlet, sub-outlet)");
}

// ---- the file ------------------------------------------------------------------

Case read_case(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error(path + ": cannot open");
    std::map<std::string, std::map<std::string, std::string>> ini;
    std::string sec, line;
    int ln = 0;
    while (std::getline(f, line)) {
        ln++;
        const size_t c = line.find_first_of(";#");
        if (c != std::string::npos) line = line.substr(0, c);
        line = trim(line);
        if (line.empty()) continue;
        if (line.front() == '[') {
            if (line.back() != ']') throw std::runtime_error(path + ":" + std::to_string(ln) + ": bad section header");
            sec = trim(line.substr(1, line.size() - 2));
            ini[sec];
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos || sec.empty())
            throw std::runtime_error(path + ":" + std::to_string(ln) + ": expected key = value inside a section");
        ini[sec][lower(trim(line.substr(0, eq)))] = trim(line.substr(eq + 1));
    }
    auto get = [&](const std::string& s, const std::string& k, const std::string& dflt, bool required) {
        auto it = ini.find(s);
        if (it != ini.end()) {
            auto jt = it->second.find(k);
            if (jt != it->second.end()) return jt->second;
        }
        if (required) throw std::runtime_error(path + ": [" + s + "] " + k + " is required");
        return dflt;
    };
    auto num = [&](const std::string& s, const std::string& k, double dflt, bool required) {
        const std::string v = get(s, k, "", required);
        if (v.empty()) return dflt;
        try {
            return Expr(v)(0.0, 0.0);
        } catch (const std::exception& e) {
            throw std::runtime_error(path + ": [" + s + "] " + k + ": " + e.what());
        }
    };
    Case c;
    const std::string dir = path.find('/') == std::string::npos ? "" : path.substr(0, path.rfind('/') + 1);
    c.mesh = get("mesh", "file", "", true);
    if (!c.mesh.empty() && c.mesh.front() != '/') c.mesh = dir + c.mesh;
    c.gamma = num("consts", "gamma", 1.4, false);
    if (!(c.gamma > 1.0)) throw std::runtime_error(path + ": [consts] gamma must exceed 1");
    const std::string prec = lower(get("solver", "precision", "fp64", false));
    if (prec == "fp64") c.precision = HCFD_FP64;
    else if (prec == "mixed") c.precision = HCFD_MIXED;
    else throw std::runtime_error(path + ": [solver] precision must be fp64 or mixed");
    const std::string order = lower(get("solver", "order", "hilbert", false));
    if (order != "hilbert" && order != "rcm") throw std::runtime_error(path + ": [solver] order must be hilbert or rcm");
    c.hilbert = order == "hilbert";
    for (const char* k : {"r", "u", "v", "p"}) {
        try {
            c.ic.emplace_back(get("ICs", k, "", true));
        } catch (const std::exception& e) {
            throw std::runtime_error(std::string(e.what()).find(path) == 0 ? e.what()
                                                                           : path + ": [ICs] " + k + ": " + e.what());
        }
    }
    for (auto& kv : ini) {
        if (kv.first.rfind("BC-", 0) != 0) continue;
        BC b;
        b.name = kv.first.substr(3);
        try {
            b.type = bc_type_of(get(kv.first, "type", "", true));
            const bool has = kv.second.count("r") || kv.second.count("u") || kv.second.count("v") || kv.second.count("p");
            if (has)
                for (const char* k : {"r", "u", "v", "p"}) b.q.emplace_back(get(kv.first, k, "", true));
        } catch (const std::exception& e) {
            throw std::runtime_error(std::string(e.what()).find(path) == 0 ? e.what()
                                                                           : path + ": [" + kv.first + "] " + e.what());
        }
        c.bcs[b.name] = b;
    }
    const std::string scheme = lower(get("time-integrator", "scheme", "euler", false));
    if (scheme == "euler") c.stages = 1;
    else if (scheme == "ssprk2") c.stages = 2;
    else if (scheme == "ssprk3") c.stages = 3;
    else throw std::runtime_error(path + ": [time-integrator] scheme must be euler, ssprk2 or ssprk3");
    c.T = num("time-integrator", "t", 0.0, true);
    c.dt = num("time-integrator", "dt", 0.0, false);
    c.cfl = num("time-integrator", "cfl", 0.0, false);
    if ((c.dt > 0) == (c.cfl > 0))
        throw std::runtime_error(path + ": [time-integrator] needs exactly one of dt > 0 and cfl > 0");
    if (!(c.T > 0)) throw std::runtime_error(path + ": [time-integrator] T must be positive");
    c.dt_out = num("time-integrator", "dt-out", c.T, false);
    if (!(c.dt_out > 0)) throw std::runtime_error(path + ": [time-integrator] dt-out must be positive");
    c.stem = get("output", "stem", "", false);
    if (c.stem.empty()) {
        const std::string base = path.substr(path.rfind('/') + 1);
        c.stem = dir + base.substr(0, base.rfind('.'));
    } else if (c.stem.front() != '/') {
        c.stem = dir + c.stem;
    }
    return c;
}

}  // namespace casefile
