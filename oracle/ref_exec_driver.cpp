// ref_exec_driver.cpp -- runs a flattened QRunes instruction list through
// QPanda's own executor and CPU simulator.
//
// TEST INFRASTRUCTURE ONLY (the oracle's reference leg for QRunes programs).
// Linked against the unmodified QPandaSDK/QuantumInstructionHandle/
// {Instruction,QuantumInstructionGet,QuantumGates,CPUQuantumGates}.cpp, compiled
// straight from the reference tree by oracle/Makefile into oracle/_ref/.
//
// The reference's parser (QRunesParser/lex.yy.cpp) and its instruction handle
// (QuantumInstructionHandleAPI.cpp, which includes <Windows.h>) do not build
// here, so this driver takes the instruction list the engine's parser prints
// (one line per instruction: "<kind> q... | angle=.. dagger=0|1 ctrl=a,b
// creg=c"), builds the reference's own node structs (QRunesParser/QParamStruct.h)
// from it, and repeats the two loops those files hold, in its own code:
//   * QuantumInstructionHandleAPI::transmit (QuantumInstructionHandleAPI.cpp:68-118):
//     an END node first (endGate: the state back to |0...0>), then every node
//     through QuantumInstructionGet::transmit until one fails;
//   * QPandaSDK::run / countState / getResult (QPandaAPI.cpp:130-305): the
//     program once, MEASURE programs repeated iRepeat times with the register
//     strings counted (registers ascending, string reversed) and divided by
//     iRepeat; PMEASURE programs reported under MSB-first bit strings.
// The generator is seeded as in ref_driver.cpp: localtime() is defined here
// and returns tm_year = seed - 70, zeros elsewhere, so the reference's seed
// expression (QuantumGates.cpp:63-72) evaluates to the seed.
//
// Usage: qpanda_ref_exec LIST SEED REPEAT RUNS OUTPREFIX
//   LIST's first line is "QINIT <n>"; RUNS consecutive runs of REPEAT shots
//   continue one generator stream (one process, as in the reference).
//   Writes OUTPREFIX.<i>.txt (the result text of run i, as getResult prints
//   it) and OUTPREFIX.state (final state: 2^n x {f64 re, f64 im}); prints
//   the QError code of every run.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "CPUQuantumGates.h"
#include "Instruction.h"
#include "QError.h"
#include "QuantumInstructionGet.h"

static struct tm g_tm;

extern "C" struct tm *localtime(const time_t *) { return &g_tm; }

namespace {

struct Parsed {
    std::string kind;
    std::vector<int> qubits;
    double angle = 0;
    bool dagger = false;
    std::vector<size_t> controls;
    int creg = 0;
};

bool parse_line(const std::string &line, Parsed &p)
{
    const size_t bar = line.find(" | ");
    if (bar == std::string::npos) return false;
    std::istringstream head(line.substr(0, bar));
    head >> p.kind;
    int q;
    while (head >> q) p.qubits.push_back(q);
    std::istringstream tail(line.substr(bar + 3));
    std::string kv;
    while (tail >> kv) {
        const size_t eq = kv.find('=');
        const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
        if (k == "angle") p.angle = std::strtod(v.c_str(), nullptr);
        else if (k == "dagger") p.dagger = v == "1";
        else if (k == "creg") p.creg = std::atoi(v.c_str());
        else if (k == "ctrl") {
            std::istringstream cs(v);
            std::string c;
            while (std::getline(cs, c, ','))
                if (!c.empty()) p.controls.push_back((size_t)std::atoll(c.c_str()));
        }
    }
    return true;
}

template <class T> T *gate(int spec, const Parsed &p)
{
    T *n = new T();
    n->gateSpecifier = spec;
    n->next = nullptr;
    n->isDagger = p.dagger;
    n->controlList = p.controls;
    return n;
}

// One reference node per instruction, the struct the executor casts it to
// (Instruction.cpp:53, 104, 145, 233, 289, 343, 398, 451, 509, 551, 609, 701, 811).
QNode *make_node(const Parsed &p)
{
    const std::string &k = p.kind;
    if (k == "H" || k == "NOT" || k == "RESET") {
        auto *n = gate<SingleGateNode>(k == "H" ? HADAMARD_GATE : k == "NOT" ? NOT_GATE : RESET_GATE, p);
        n->opQubit = p.qubits.at(0);
        return n;
    }
    if (k == "RX" || k == "RY" || k == "RZ") {
        auto *n = gate<SingleAngleGateNode>(k == "RX" ? RX_GATE : k == "RY" ? RY_GATE : RZ_GATE, p);
        n->opQubit = p.qubits.at(0);
        n->angle = p.angle;
        return n;
    }
    if (k == "CNOT" || k == "SQISWAP") {
        auto *n = gate<DoubleGateNode>(k == "CNOT" ? CNOT_GATE : SQISWAP_GATE, p);
        n->opQubit1 = p.qubits.at(0);
        n->opQubit2 = p.qubits.at(1);
        return n;
    }
    if (k == "CR" || k == "ISWAP") {
        auto *n = gate<DoubleAngleGateNode>(k == "CR" ? CR_GATE : ISWAP_GATE, p);
        n->opQubit1 = p.qubits.at(0);
        n->opQubit2 = p.qubits.at(1);
        n->angle = p.angle;
        return n;
    }
    if (k == "TOFFOLI") {
        auto *n = gate<TripleGateNode>(TOFFOLI_GATE, p);
        n->opQubit1 = p.qubits.at(0);
        n->opQubit2 = p.qubits.at(1);
        n->opQubit3 = p.qubits.at(2);
        return n;
    }
    if (k == "MEASURE") {
        auto *n = new MeasureNode();
        n->gateSpecifier = MEASURE_GATE;
        n->next = nullptr;
        n->opQubit = p.qubits.at(0);
        n->creg = p.creg;
        return n;
    }
    if (k == "PMEASURE") {
        auto *n = new pMeasureNode();
        n->gateSpecifier = PMEASURE_GATE;
        n->next = nullptr;
        for (int q : p.qubits) n->vecint_pmeasure_list.push_back((size_t)q);
        return n;
    }
    return nullptr;
}

bool transmit(const std::vector<QNode *> &nodes, QuantumGateParam *param, QuantumGates *g)
{
    QuantumInstructionGet *get = QuantumInstructionGet::getInstance();
    if (g->getQStateSize() == 0) return false;
    QNode endProg = {END, nullptr};
    if (!get->transmit(&endProg, param, g)) return false;
    for (QNode *n : nodes)
        if (!get->transmit(n, param, g)) return false;
    return true;
}

std::string count_key(const QuantumGateParam &param)
{
    std::stringstream ss;
    for (auto &kv : param.mReturnValue) ss << kv.second;
    std::string s(ss.str());
    return std::string(s.rbegin(), s.rend());
}

std::string msb_first(size_t number, size_t len)
{
    std::string s;
    unsigned int index = 1u << (len - 1);
    for (size_t i = 0; i < len; i++, index >>= 1) s += (number & index) ? '1' : '0';
    return s;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s LIST SEED REPEAT RUNS OUTPREFIX\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    std::string line;
    int n = 0;
    if (!std::getline(in, line) || std::sscanf(line.c_str(), "QINIT %d", &n) != 1 || n < 1 || n > 30) {
        std::fprintf(stderr, "bad LIST header\n");
        return 2;
    }
    std::vector<QNode *> nodes;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        Parsed p;
        QNode *node = parse_line(line, p) ? make_node(p) : nullptr;
        if (!node) {
            std::fprintf(stderr, "bad instruction: %s\n", line.c_str());
            return 2;
        }
        if (!nodes.empty()) nodes.back()->next = node;
        nodes.push_back(node);
    }
    g_tm = {};
    g_tm.tm_year = std::atoi(argv[2]) - 70;
    const int repeat = std::atoi(argv[3]), runs = std::atoi(argv[4]);
    const std::string prefix = argv[5];

    QuantumGateParam param;
    param.setQBitNum(n);
    CPUQuantumGates g;
    bool is_run = false;
    for (int r = 0; r < runs; r++) {
        param.mPMeasure.clear();
        param.mReturnValue.clear();
        std::map<std::string, double> result;
        int code = qErrorNone;
        if (!is_run && g.initState(&param) != qErrorNone) code = initStateError;
        if (code == qErrorNone) {
            if (!transmit(nodes, &param, &g)) {
                code = runProgramError;
                is_run = false;
            }
            if (param.mPMeasureSize == 0 && code == qErrorNone) {
                result[count_key(param)] += 1;
                for (int i = 1; i < repeat; i++) {
                    if (!transmit(nodes, &param, &g)) {
                        code = runProgramError;
                        is_run = false;
                        break;
                    }
                    result[count_key(param)] += 1;
                }
                for (auto &kv : result) kv.second /= repeat;
            } else {
                for (auto &pm : param.mPMeasure) result.insert({msb_first(pm.first, param.mPMeasureSize), pm.second});
            }
            if (code == qErrorNone) is_run = true;
        }
        std::ostringstream text;
        if (is_run)
            for (auto &kv : result) text << kv.first << " " << kv.second << "\n";
        std::ofstream(prefix + "." + std::to_string(r) + ".txt") << text.str();
        std::printf("%d\n", code);
    }
    // the final state from the simulator's amplitude vector (getQState's
    // text form rounds to 6 digits)
    std::ofstream(prefix + ".state", std::ios::binary)
        .write(reinterpret_cast<const char *>(g.mvQuantumStat.data()),
               (std::streamsize)(g.mvQuantumStat.size() * 16));
    // the nodes stay allocated until the process ends (the node structs have
    // no virtual destructor)
    return 0;
}
