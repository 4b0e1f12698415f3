// QRunes.cpp -- QRunes lexer, parser and DAGGER / CONTROL list controller.
// The accepted language and the list semantics are the reference front end's
// (see QRunes.h); the implementation is a hand-written scanner and a
// recursive-descent parser over vectors.
#include "QRunes.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace qsvq {

namespace {

enum Tok {
    T_EOF,
    T_QINIT,
    T_CREG,
    T_DAGGER,
    T_ENDDAGGER,
    T_INT,
    T_FLOAT,
    T_SINGLE,
    T_DOUBLE,
    T_TRIPLE,
    T_SINGLEANGLE,
    T_DOUBLEANGLE,
    T_MEASURE,
    T_PMEASURE,
    T_CONTROL,
    T_ENDCONTROL,
    T_CHAR          // one of + - * / ^ ( ) " $ ,
};

struct Token {
    Tok t = T_EOF;
    int gate = 0;
    long ival = 0;
    double dval = 0;
    char ch = 0;
    int line = 1;
};

struct Keyword {
    const char *text;
    Tok t;
    int gate;
};

// lex.yy.cpp:914-1079 (rules 4-27)
const Keyword kKeywords[] = {
    {"QINIT", T_QINIT, 0},
    {"CREG", T_CREG, 0},
    {"DAGGER", T_DAGGER, 0},
    {"ENDDAGGER", T_ENDDAGGER, 0},
    {"HADAMARD", T_SINGLE, HADAMARD_GATE},
    {"H", T_SINGLE, HADAMARD_GATE},
    {"NOT", T_SINGLE, NOT_GATE},
    {"RESET", T_SINGLE, RESET_GATE},
    {"RX", T_SINGLEANGLE, RX_GATE},
    {"RY", T_SINGLEANGLE, RY_GATE},
    {"RZ", T_SINGLEANGLE, RZ_GATE},
    {"ISWAP", T_DOUBLEANGLE, ISWAP_GATE},
    {"CR", T_DOUBLEANGLE, CR_GATE},
    {"CPHASE", T_DOUBLEANGLE, CR_GATE},
    {"SQISWAP", T_DOUBLE, SQISWAP_GATE},
    {"CNOT", T_DOUBLE, CNOT_GATE},
    {"TOFFOLI", T_TRIPLE, TOFFOLI_GATE},
    {"MEASURE", T_MEASURE, MEASURE_GATE},
    {"PMEASURE", T_PMEASURE, PMEASURE_GATE},
    {"CONTROL", T_CONTROL, 0},
    {"ENDCONTROL", T_ENDCONTROL, 0},
};

const double kPi = 3.14159265358979323846;   // lex.yy.cpp:939

class Lexer {
public:
    explicit Lexer(const std::string &s) : s_(s) {}

    Token next()
    {
        for (;;) {
            if (p_ >= s_.size()) {
                Token t;
                t.line = line_;
                return t;
            }
            const char c = s_[p_];
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f') {
                if (c == '\n') line_++;
                p_++;
                continue;
            }
            if (c == '%') {                        // line comment, newline included
                while (p_ < s_.size() && s_[p_] != '\n') p_++;
                if (p_ < s_.size()) {
                    p_++;
                    line_++;
                }
                continue;
            }
            // longest match over the token rules
            size_t best = 0;
            Token tok;
            tok.line = line_;
            for (const Keyword &k : kKeywords) {
                const size_t n = std::char_traits<char>::length(k.text);
                if (n > best && s_.compare(p_, n, k.text) == 0) {
                    best = n;
                    tok.t = k.t;
                    tok.gate = k.gate;
                }
            }
            if (p_ + 2 <= s_.size() && std::tolower((unsigned char)s_[p_]) == 'p' &&
                std::tolower((unsigned char)s_[p_ + 1]) == 'i' && best < 2) {
                best = 2;
                tok.t = T_FLOAT;
                tok.dval = kPi;
            }
            size_t d = 0;
            while (p_ + d < s_.size() && std::isdigit((unsigned char)s_[p_ + d])) d++;
            if (d > 0) {
                size_t len = d;
                Tok t = T_INT;
                if (p_ + d + 1 < s_.size() && s_[p_ + d] == '.' && std::isdigit((unsigned char)s_[p_ + d + 1])) {
                    size_t f = d + 1;
                    while (p_ + f < s_.size() && std::isdigit((unsigned char)s_[p_ + f])) f++;
                    len = f;
                    t = T_FLOAT;
                }
                if (len > best) {
                    best = len;
                    tok.t = t;
                    const std::string lit = s_.substr(p_, len);
                    // rule 8 reads integers with atoi, (int) strtol: beyond INT_MAX the
                    // value keeps its low 32 bits, as there
                    if (t == T_INT) tok.ival = (int)std::strtol(lit.c_str(), nullptr, 10);
                    else tok.dval = std::strtod(lit.c_str(), nullptr);
                }
            }
            if (best == 0 && std::string("+-*/^()\"$,").find(c) != std::string::npos) {
                best = 1;
                tok.t = T_CHAR;
                tok.ch = c;
            }
            if (best == 0) {                      // rule 40: any other character is echoed, no token
                p_++;
                continue;
            }
            p_ += best;
            return tok;
        }
    }

private:
    const std::string &s_;
    size_t p_ = 0;
    int line_ = 1;
};

// The list controller (LinkedListController.cpp:51-339) over vectors.
class ListController {
public:
    std::vector<QRNode> root;

    void gate(QRNode n)
    {
        n.controls = controls_;
        if (stack_.empty()) {
            root.push_back(std::move(n));
        } else {
            n.dagger = stack_.back().level % 2 == 0;
            stack_.back().nodes.push_back(std::move(n));
        }
    }
    void measure(QRNode n)                           // also PMEASURE
    {
        if (!stack_.empty() || !controls_.empty()) return;
        root.push_back(std::move(n));
    }
    void daggerOpen()
    {
        const size_t parent = stack_.empty() ? 1 : stack_.back().level;
        stack_.push_back({{}, parent + 1});
    }
    void daggerClose()
    {
        if (stack_.empty()) return;
        Sub top = std::move(stack_.back());
        stack_.pop_back();
        if (stack_.empty()) {
            std::reverse(top.nodes.begin(), top.nodes.end());
            root.insert(root.end(), top.nodes.begin(), top.nodes.end());
        } else {
            if (top.level % 2 == 1) std::reverse(top.nodes.begin(), top.nodes.end());
            auto &dst = stack_.back().nodes;
            dst.insert(dst.end(), top.nodes.begin(), top.nodes.end());
        }
    }
    void controlOpen(size_t q)
    {
        if (std::count(controls_.begin(), controls_.end(), q) == 0) controls_.push_back(q);
    }
    void controlClose(size_t q)
    {
        auto it = std::find(controls_.begin(), controls_.end(), q);
        if (it != controls_.end()) controls_.erase(it);
    }

private:
    struct Sub {
        std::vector<QRNode> nodes;
        size_t level;
    };
    std::vector<Sub> stack_;
    std::vector<size_t> controls_;
};

class Parser {
public:
    Parser(const std::string &text, QRProgram &prog) : lex_(text), prog_(prog) { advance(); }

    bool run(std::string *error)
    {
        const bool ok = file();
        if (!ok && error) *error = err_;
        return ok;
    }

private:
    Lexer lex_;
    QRProgram &prog_;
    Token cur_;
    ListController list_;
    std::string err_;
    long qnum_ = 0, cnum_ = 0;

    void advance() { cur_ = lex_.next(); }
    bool fail(const std::string &m)
    {
        if (err_.empty()) err_ = "line " + std::to_string(cur_.line) + ": " + m;
        return false;
    }
    bool isChar(char c) const { return cur_.t == T_CHAR && cur_.ch == c; }
    bool expectChar(char c)
    {
        if (!isChar(c)) return fail(std::string("syntax error, expected '") + c + "'");
        advance();
        return true;
    }
    // an integer in an index position (QINIT, CREG, qubits, classical bits);
    // one that wrapped to a negative value is refused, where the reference
    // would go on to use it as an index
    bool expectInt(long *v)
    {
        if (cur_.t != T_INT) return fail("syntax error, expected an integer");
        if (cur_.ival < 0) return fail("integer out of range");
        *v = cur_.ival;
        advance();
        return true;
    }
    bool qubit(long q) { return q < qnum_ ? true : fail("qubit out of bounds"); }

    // file: qinit cinit gatelist [pmeasure]   (y.tab.cpp rules 2-3, 7-10)
    bool file()
    {
        if (cur_.t != T_QINIT) return fail("syntax error, expected QINIT");
        advance();
        if (!expectInt(&qnum_)) return false;
        if (cur_.t != T_CREG) return fail("syntax error, expected CREG");
        advance();
        if (!expectInt(&cnum_)) return false;
        int lines = 0;
        while (startsLine()) {
            if (!gateline()) return false;
            lines++;
        }
        if (lines == 0) return fail("syntax error, expected a statement");
        if (cur_.t == T_PMEASURE && !pmeasure()) return false;
        if (cur_.t != T_EOF) return fail("syntax error");
        prog_.qubits = (int)qnum_;
        prog_.cregs = (int)cnum_;
        prog_.nodes = std::move(list_.root);
        return true;
    }

    bool startsLine() const
    {
        switch (cur_.t) {
        case T_MEASURE: case T_CONTROL: case T_ENDCONTROL: case T_DAGGER: case T_ENDDAGGER:
        case T_SINGLE: case T_DOUBLE: case T_TRIPLE: case T_SINGLEANGLE: case T_DOUBLEANGLE:
            return true;
        default:
            return false;
        }
    }

    // y.tab.cpp rules 11-20
    bool gateline()
    {
        const Token t = cur_;
        advance();
        long a = 0, b = 0, c = 0;
        QRNode n;
        n.kind = t.gate;
        switch (t.t) {
        case T_MEASURE:
            if (!expectInt(&a) || !expectChar(',') || !expectChar('$') || !expectInt(&b)) return false;
            if (!qubit(a)) return false;
            if (b >= cnum_) return fail("classical register out of bounds");
            n.kind = MEASURE_GATE;
            n.qubits = {(size_t)a};
            n.creg = (size_t)b;
            list_.measure(std::move(n));
            return true;
        case T_CONTROL:
            if (!expectInt(&a) || !qubit(a)) return false;
            list_.controlOpen((size_t)a);
            return true;
        case T_ENDCONTROL:
            if (!expectInt(&a) || !qubit(a)) return false;
            list_.controlClose((size_t)a);
            return true;
        case T_DAGGER:
            list_.daggerOpen();
            return true;
        case T_ENDDAGGER:
            list_.daggerClose();
            return true;
        case T_SINGLE:
            if (!expectInt(&a) || !qubit(a)) return false;
            n.qubits = {(size_t)a};
            break;
        case T_DOUBLE:
            if (!expectInt(&a) || !expectChar(',') || !expectInt(&b)) return false;
            if (!qubit(a) || !qubit(b)) return false;
            n.qubits = {(size_t)a, (size_t)b};
            break;
        case T_TRIPLE:
            if (!expectInt(&a) || !expectChar(',') || !expectInt(&b) || !expectChar(',') || !expectInt(&c))
                return false;
            if (!qubit(a) || !qubit(b) || !qubit(c)) return false;
            n.qubits = {(size_t)a, (size_t)b, (size_t)c};
            break;
        case T_SINGLEANGLE:
            if (!expectInt(&a) || !expectChar(',') || !expectChar('"') || !additive(&n.angle) || !expectChar('"'))
                return false;
            if (!qubit(a)) return false;
            n.qubits = {(size_t)a};
            break;
        case T_DOUBLEANGLE:
            if (!expectInt(&a) || !expectChar(',') || !expectInt(&b) || !expectChar(',') || !expectChar('"') ||
                !additive(&n.angle) || !expectChar('"'))
                return false;
            if (!qubit(a) || !qubit(b)) return false;
            n.qubits = {(size_t)a, (size_t)b};
            break;
        default:
            return fail("syntax error");
        }
        list_.gate(std::move(n));
        return true;
    }

    // y.tab.cpp rules 4-6
    bool pmeasure()
    {
        advance();
        QRNode n;
        n.kind = PMEASURE_GATE;
        long q = 0;
        if (!expectInt(&q) || !qubit(q)) return false;
        n.qubits.push_back((size_t)q);
        while (isChar(',')) {
            advance();
            if (!expectInt(&q) || !qubit(q)) return false;
            n.qubits.push_back((size_t)q);
        }
        list_.measure(std::move(n));
        return true;
    }

    // y.tab.cpp rules 21-34
    bool additive(double *v)
    {
        if (!multiplicative(v)) return false;
        while (isChar('+') || isChar('-')) {
            advance();
            double r;
            if (!multiplicative(&r)) return false;
            *v = *v + r;                               // rule 24 adds for '-' as well
        }
        return true;
    }
    bool multiplicative(double *v)
    {
        if (!power(v)) return false;
        while (isChar('*') || isChar('/')) {
            const bool div = isChar('/');
            advance();
            double r;
            if (!power(&r)) return false;
            if (div) {
                *v = *v / r;
                if (!std::isnormal(*v)) return fail("angle error");
            } else {
                *v = *v * r;
            }
        }
        return true;
    }
    bool power(double *v)
    {
        if (!primary(v)) return false;
        while (isChar('^')) {
            advance();
            double r;
            if (!primary(&r)) return false;
            *v = std::pow(*v, r);
            if (!std::isnormal(*v)) return fail("angle error");
        }
        return true;
    }
    bool primary(double *v)
    {
        if (isChar('(')) {
            advance();
            return additive(v) && expectChar(')');
        }
        if (isChar('-') || isChar('+')) {
            const bool neg = isChar('-');
            advance();
            if (!primary(v)) return false;
            if (neg) *v = -*v;
            return true;
        }
        if (cur_.t == T_INT) {
            *v = (double)cur_.ival;
            advance();
            return true;
        }
        if (cur_.t == T_FLOAT) {
            *v = cur_.dval;
            advance();
            return true;
        }
        return fail("syntax error in angle");
    }
};

const char *kindName(int k)
{
    switch (k) {
    case HADAMARD_GATE: return "H";
    case RX_GATE: return "RX";
    case RY_GATE: return "RY";
    case RZ_GATE: return "RZ";
    case NOT_GATE: return "NOT";
    case TOFFOLI_GATE: return "TOFFOLI";
    case RESET_GATE: return "RESET";
    case CR_GATE: return "CR";
    case ISWAP_GATE: return "ISWAP";
    case SQISWAP_GATE: return "SQISWAP";
    case CNOT_GATE: return "CNOT";
    case MEASURE_GATE: return "MEASURE";
    case PMEASURE_GATE: return "PMEASURE";
    default: return "?";
    }
}

}  // namespace

int parseQRunesText(const std::string &text, QRProgram &prog, std::string *error)
{
    QRProgram p;
    Parser parser(text, p);
    if (!parser.run(error)) return -1;
    prog = std::move(p);
    return prog.qubits;
}

int parseQRunesFile(const std::string &path, QRProgram &prog, std::string *error)
{
    if (path.empty()) {
        if (error) *error = "empty path";
        return -1;
    }
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        if (error) *error = "cannot open " + path;
        return -1;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parseQRunesText(ss.str(), prog, error);
}

std::string describe(const QRProgram &prog)
{
    std::string out;
    char buf[64];
    for (const QRNode &n : prog.nodes) {
        out += kindName(n.kind);
        for (size_t q : n.qubits) out += " " + std::to_string(q);
        std::snprintf(buf, sizeof buf, " | angle=%.17g dagger=%d ctrl=", n.angle, n.dagger ? 1 : 0);
        out += buf;
        for (size_t i = 0; i < n.controls.size(); i++) out += (i ? "," : "") + std::to_string(n.controls[i]);
        out += " creg=" + std::to_string(n.creg) + "\n";
    }
    return out;
}

}  // namespace qsvq
