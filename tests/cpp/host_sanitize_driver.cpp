// host_sanitize_driver.cpp -- drives the index-heavy host code under
// AddressSanitizer / UndefinedBehaviorSanitizer (built by
// tests/test_host_sanitizers.py): the greedy block planner (qsv_fuse_plan)
// and the window planners (qsv_window_plan, qsv_window_plan_beam) on random gate lists of every
// record kind, and the QRunes front end on the reference's example programs,
// generated programs and byte-level mutations of them.  Any sanitizer report
// aborts the run; a planner error on a valid list is a failure too.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "QRunes.h"
#include "qsv.h"

namespace qsv {
std::vector<std::vector<std::pair<int, int>>> qubit_routes(std::vector<int> dest, int tb);
}

namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    }
    int below(int n) { return (int)(next() % (uint64_t)n); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

int g_failures = 0;

void check(bool ok, const char *what)
{
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s (%s)\n", what, qsv_last_error());
        g_failures++;
    }
}

// Random valid record list for an n-qubit state; mats receives the U1 / U2
// matrices (complex doubles).
std::vector<qsv_gate> random_gates(Rng &r, int n, int count, std::vector<double> &mats)
{
    std::vector<qsv_gate> g;
    auto pick = [&](int k, int *out) {
        for (int i = 0; i < k; i++) {
            int q;
            bool dup;
            do {
                q = r.below(n);
                dup = false;
                for (int j = 0; j < i; j++) dup |= out[j] == q;
            } while (dup);
            out[i] = q;
        }
    };
    for (int i = 0; i < count; i++) {
        qsv_gate x{};
        int maxk = n >= 3 ? 16 : (n == 2 ? 11 : 6);
        int kinds[16] = {QSV_GATE_H, QSV_GATE_X, QSV_GATE_RX, QSV_GATE_RY, QSV_GATE_RZ, QSV_GATE_U1,
                         QSV_GATE_MEASURE, QSV_GATE_RESET, QSV_GATE_CNOT, QSV_GATE_CZ, QSV_GATE_CR, QSV_GATE_ISWAP,
                         QSV_GATE_SQISWAP, QSV_GATE_U2, QSV_GATE_TOFFOLI, QSV_GATE_CSWAP};
        x.kind = kinds[r.below(maxk)];
        x.dagger = r.below(2);
        x.theta = 6.283185307179586 * r.uniform();
        x.mat = -1;
        int q[3] = {0, 0, 0};
        switch (x.kind) {
        case QSV_GATE_CNOT: case QSV_GATE_CZ: case QSV_GATE_CR: case QSV_GATE_ISWAP: case QSV_GATE_SQISWAP:
        case QSV_GATE_U2:
            pick(2, q);
            if (x.kind == QSV_GATE_CNOT && r.below(20) == 0) q[1] = q[0];   // the reference's no-op form
            break;
        case QSV_GATE_TOFFOLI: case QSV_GATE_CSWAP: pick(3, q); break;
        default: pick(1, q); break;
        }
        x.q0 = q[0];
        x.q1 = q[1];
        x.q2 = q[2];
        if (x.kind == QSV_GATE_MEASURE) x.q1 = r.below(64);
        if (x.kind == QSV_GATE_U1 || x.kind == QSV_GATE_U2) {
            const int d = x.kind == QSV_GATE_U1 ? 2 : 4;
            x.mat = (int)(mats.size() / 2);
            for (int e = 0; e < d * d; e++) {
                mats.push_back(r.uniform() - 0.5);
                mats.push_back(r.uniform() - 0.5);
            }
        }
        g.push_back(x);
    }
    return g;
}

void planners(Rng &r, int rounds)
{
    for (int it = 0; it < rounds; it++) {
        const int n = 1 + r.below(16);
        const int count = r.below(5) == 0 ? 0 : 1 + r.below(300);
        std::vector<double> mats;
        std::vector<qsv_gate> g = random_gates(r, n, count, mats);
        const int64_t cap = count + 1;
        std::vector<int> kq(cap), qs(5 * cap), win(cap);
        std::vector<double> m(2048 * cap);
        std::vector<int64_t> ng(cap);
        for (int k = 1; k <= 5; k++) {
            int64_t nb = 0;
            const int rc = qsv_fuse_plan(g.data(), count, mats.empty() ? nullptr : mats.data(), k, r.below(2), cap,
                                         kq.data(), qs.data(), m.data(), ng.data(), &nb);
            check(rc == QSV_OK, "qsv_fuse_plan");
            int64_t total = 0;
            for (int64_t b = 0; b < nb && rc == QSV_OK; b++) total += kq[b] ? ng[b] : 1;
            if (rc == QSV_OK && total != count)
                std::fprintf(stderr, "n %d k %d count %d total %lld blocks %lld\n", n, k, count, (long long)total,
                             (long long)nb);
            check(rc != QSV_OK || total == count, "qsv_fuse_plan gate count");
        }
        for (int k = 4; k <= 5; k++)
            for (int tb = 4; tb <= 13; tb++) {
                int64_t nb = 0;
                const int rc = qsv_window_plan(g.data(), count, mats.empty() ? nullptr : mats.data(), k, r.below(2),
                                               tb, n, cap, kq.data(), qs.data(), m.data(), ng.data(), win.data(), &nb);
                if (n < 4 || (tb < 7 && tb < n)) {
                    check(rc == QSV_ERR_ARG, "qsv_window_plan rejects fewer than 4 qubits and tile bits below 7");
                    continue;
                }
                check(rc == QSV_OK, "qsv_window_plan");
                int64_t total = 0;
                for (int64_t b = 0; b < nb && rc == QSV_OK; b++) total += kq[b] ? ng[b] : 1;
                check(rc != QSV_OK || total == count, "qsv_window_plan gate count");
                // the beam search over windows (the learned schedule's planner)
                const int width = 1 + r.below(4), branch = 1 + r.below(3), first = r.below(7), delay = r.below(5);
                const int rb = qsv_window_plan_beam(g.data(), count, mats.empty() ? nullptr : mats.data(), k, 0, tb,
                                                    n, width, branch, first, delay, r.below(9), r.below(2), cap, kq.data(),
                                                    qs.data(), m.data(), ng.data(), win.data(), &nb);
                check(rb == QSV_OK, "qsv_window_plan_beam");
                total = 0;
                for (int64_t b = 0; b < nb && rb == QSV_OK; b++) total += kq[b] ? ng[b] : 1;
                check(rb != QSV_OK || total == count, "qsv_window_plan_beam gate count");
            }
    }
}

std::string generated_program(Rng &r)
{
    const int n = 1 + r.below(12), cr = 1 + r.below(6);
    std::ostringstream o;
    o << "QINIT " << n << "\nCREG " << cr << "\n";
    const char *gates[] = {"H", "HADAMARD", "NOT", "RX", "RY", "RZ", "CNOT", "CR", "CPHASE", "ISWAP", "SQISWAP",
                           "TOFFOLI", "RESET", "MEASURE", "DAGGER", "ENDDAGGER", "CONTROL", "ENDCONTROL"};
    const char *angles[] = {"\"pi/3\"", "\"-pi/7\"", "\"0.25\"", "\"2*pi/5\"", "\"1-0.5\"", "\"pi^0.5\"", "\"(pi+1)/3\"",
                            "\"1/0\"", "\"0^-1\"", "\"((pi)\""};
    const int len = r.below(60);
    for (int i = 0; i < len; i++) {
        const char *gname = gates[r.below(18)];
        std::string s = gname;
        const int q0 = r.below(n + 1), q1 = r.below(n + 1), q2 = r.below(n + 1);
        if (s == "CNOT" || s == "SQISWAP") o << s << " " << q0 << "," << q1 << "\n";
        else if (s == "CR" || s == "CPHASE" || s == "ISWAP") o << s << " " << q0 << "," << q1 << "," << angles[r.below(10)] << "\n";
        else if (s == "TOFFOLI") o << s << " " << q0 << "," << q1 << "," << q2 << "\n";
        else if (s == "RX" || s == "RY" || s == "RZ") o << s << " " << q0 << "," << angles[r.below(10)] << "\n";
        else if (s == "MEASURE") o << s << " " << q0 << ",$" << r.below(cr + 1) << "\n";
        else if (s == "DAGGER" || s == "ENDDAGGER") o << s << "\n";
        else o << s << " " << q0 << "\n";
    }
    if (r.below(3) == 0) {
        o << "PMEASURE 0";
        for (int q = 1; q < n; q++) o << "," << q;
        o << "\n";
    }
    return o.str();
}

std::string mutate(Rng &r, std::string s)
{
    static const char alphabet[] = "QINITCREGHXRYZMASUPDOL0123456789,.$\"()+-*/^% \n\t\x01\xff";
    const int alen = (int)sizeof(alphabet) - 1;
    const int muts = 1 + r.below(6);
    for (int i = 0; i < muts && !s.empty(); i++) {
        const size_t pos = (size_t)r.below((int)s.size());
        switch (r.below(4)) {
        case 0: s.erase(pos, (size_t)r.below(8)); break;
        case 1: s.insert(pos, 1, alphabet[r.below(alen)]); break;
        case 2: s[pos] = alphabet[r.below(alen)]; break;
        default: s = s.substr(0, pos); break;
        }
    }
    return s;
}

void qrunes(Rng &r, const std::vector<std::string> &files, int rounds)
{
    std::vector<std::string> corpus;
    for (const std::string &f : files) {
        std::ifstream in(f);
        std::stringstream ss;
        ss << in.rdbuf();
        corpus.push_back(ss.str());
        qsvq::QRProgram p;
        if (qsvq::parseQRunesFile(f, p) >= 0) (void)qsvq::describe(p);
    }
    for (int it = 0; it < rounds; it++) {
        std::string text = (!corpus.empty() && r.below(2)) ? corpus[(size_t)r.below((int)corpus.size())]
                                                           : generated_program(r);
        if (r.below(4)) text = mutate(r, text);
        qsvq::QRProgram p;
        std::string err;
        if (qsvq::parseQRunesText(text, p, &err) >= 0) (void)qsvq::describe(p);
    }
}

}  // namespace

// qubit_routes: random permutations of 4..40 positions; every pass must move
// a permutation of at most 13 positions that holds 0..3, and the passes must
// take every position's data to its destination, within the stated bound.
void routes(Rng &r, int rounds)
{
    for (int t = 0; t < rounds * 5; t++) {
        const int n = 4 + r.below(37);
        std::vector<int> dest(n);
        for (int i = 0; i < n; i++) dest[i] = i;
        const int swaps = r.below(2 * n);
        for (int i = 0; i < swaps; i++) std::swap(dest[r.below(n)], dest[r.below(n)]);
        auto passes = qsv::qubit_routes(dest, 13);
        std::vector<int> data(n);      // data[p] = the original position of the data now at p
        for (int i = 0; i < n; i++) data[i] = i;
        bool ok = true;
        for (const auto &moves : passes) {
            uint64_t from = 0, to = 0;
            for (const auto &m : moves) {
                ok &= m.first >= 0 && m.first < n && m.second >= 0 && m.second < n && m.first != m.second;
                if (!ok) break;
                ok &= !(from >> m.first & 1) && !(to >> m.second & 1);
                from |= uint64_t(1) << m.first;
                to |= uint64_t(1) << m.second;
            }
            ok &= from == to && __builtin_popcountll(from | 15) <= 13 && !moves.empty();
            if (!ok) break;
            std::vector<int> nd(data);
            for (const auto &m : moves) nd[m.second] = data[m.first];
            data.swap(nd);
        }
        for (int p = 0; ok && p < n; p++) ok &= dest[data[p]] == p;
        // cycles: ceil((c - 1) / 12) passes for the longest cycle outside 0..3 at least,
        // and never more than the cycles' sum of ceil(c / 9) + 1
        int bound = 1;
        std::vector<char> seen(n);
        for (int p = 0; p < n; p++) {
            if (seen[p]) continue;
            int c = 0;
            for (int q = p; !seen[q]; q = dest[q]) seen[q] = 1, c++;
            if (c > 1) bound += (c + 8) / 9;
        }
        ok &= (int)passes.size() <= bound;
        check(ok, "qubit routes");
    }
}

int main(int argc, char **argv)
{
    Rng r{0x9E3779B97F4A7C15ull};
    const int rounds = argc > 1 ? std::atoi(argv[1]) : 200;
    std::vector<std::string> files;
    for (int i = 2; i < argc; i++) files.push_back(argv[i]);
    planners(r, rounds);
    routes(r, rounds);
    qrunes(r, files, rounds * 20);
    std::printf("host sanitize driver: %d planner rounds, %d QRunes texts, %d failures\n", rounds, rounds * 20,
                g_failures);
    return g_failures ? 1 : 0;
}
