// ref_driver.cpp -- runs QPanda's own CPU simulator on a circuit file.
//
// TEST INFRASTRUCTURE ONLY (the oracle's reference leg).  Linked against the
// unmodified QPandaSDK/QuantumInstructionHandle/{QuantumGates,CPUQuantumGates}.cpp
// compiled straight from the reference tree by oracle/Makefile; built into
// oracle/_ref/, never into the engine.
//
// It drives CPUQuantumGates the way the reference executor does
// (Instruction.cpp:40-827): error_rate 0, controls-then-target lists, one
// program re-simulation per shot after endGate (QPandaAPI.cpp:173-182,
// QuantumInstructionHandleAPI.cpp:96-114).
//
// The reference seeds its generator once from localtime() (QuantumGates.cpp:57-72).
// This executable defines localtime() itself, returning tm_year = seed - 70 and
// zeros elsewhere, which makes the reference's seed expression evaluate to
// `seed`; calls from the reference objects bind to this definition at link time.
//
// Commands:
//   run  CIRCUIT OUTPREFIX SEED SHOTS   -> OUTPREFIX.state  (2^n x {f64 re, f64 im},
//                                          state before the first MEASURE of shot 0)
//                                          OUTPREFIX.shots  (SHOTS x u64, bit c = creg c)
//   time CIRCUIT STRIDE MAX_SECONDS     -> one JSON line: wall time of records 0, STRIDE,
//                                          2*STRIDE, ... until MAX_SECONDS has passed
//                                          (CZ as CR(pi); U2 as a 3-CNOT template)
//   check CIRCUIT NRECORDS              -> one JSON line: FNV-1a digest of the state bytes
//                                          after records 0 .. NRECORDS-1 (same substitutions),
//                                          to compare an OpenMP build with the 1-thread one
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "CPUQuantumGates.h"
#include "../include/qsv.h"

static struct tm g_tm;

extern "C" struct tm *localtime(const time_t *) { return &g_tm; }

struct Circuit {
    int n = 0;
    std::vector<qsv_gate> gates;
    std::vector<double> mats;
};

static bool load_circuit(const char *path, Circuit &c)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char magic[4];
    int32_t n;
    int64_t ng, nm;
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "QSVC", 4) == 0 &&
              std::fread(&n, 4, 1, f) == 1 && std::fread(&ng, 8, 1, f) == 1 &&
              std::fread(&nm, 8, 1, f) == 1;
    if (ok) {
        c.n = n;
        c.gates.resize((size_t)ng);
        c.mats.resize((size_t)nm * 2);
        ok = std::fread(c.gates.data(), sizeof(qsv_gate), (size_t)ng, f) == (size_t)ng &&
             std::fread(c.mats.data(), 16, (size_t)nm, f) == (size_t)nm;
    }
    std::fclose(f);
    return ok;
}

static const double kPi = 3.14159265358979323846;

// One executor-style call per record.  Returns false for records the
// reference cannot express.
static bool apply(CPUQuantumGates &g, const qsv_gate &r, std::map<size_t, int> &creg)
{
    switch (r.kind) {
    case QSV_GATE_H: g.Hadamard(r.q0, 0); return true;
    case QSV_GATE_X: g.NOT(r.q0, 0); return true;
    case QSV_GATE_RX: g.RX(r.q0, r.theta, 0, r.dagger); return true;
    case QSV_GATE_RY: g.RY(r.q0, r.theta, 0, r.dagger); return true;
    case QSV_GATE_RZ: g.RZ(r.q0, r.theta, 0, r.dagger); return true;
    case QSV_GATE_CNOT: g.CNOT(r.q0, r.q1, 0); return true;
    case QSV_GATE_CZ: g.CR(r.q0, r.q1, kPi, 0, 0); return true;
    case QSV_GATE_CR: g.CR(r.q0, r.q1, r.theta, 0, r.dagger); return true;
    case QSV_GATE_ISWAP: g.iSWAP(r.q0, r.q1, 0, r.dagger); return true;
    case QSV_GATE_SQISWAP: g.sqiSWAP(r.q0, r.q1, 0, r.dagger); return true;
    case QSV_GATE_TOFFOLI: g.toffoli(r.q0, r.q1, r.q2, 0, 0); return true;
    case QSV_GATE_CSWAP: g.controlSwap(r.q0, r.q1, r.q2, 0); return true;
    case QSV_GATE_MEASURE: creg[(size_t)r.q1] = g.qubitMeasure(r.q0); return true;
    case QSV_GATE_RESET: g.Reset(r.q0); return true;
    default: return false;
    }
}

// A Haar two-qubit gate costs the reference a KAK-shaped sequence: three
// single-qubit Euler rotations per side around three CNOTs.  Only the shape
// matters for timing (sweep cost does not depend on the angles).
static void apply_u2_template(CPUQuantumGates &g, int a, int b)
{
    for (int q : {a, b}) { g.RZ(q, 0.3, 0, 0); g.RY(q, 0.5, 0, 0); g.RZ(q, 0.7, 0, 0); }
    g.CNOT(a, b, 0);
    g.RZ(a, 0.2, 0, 0); g.RY(b, 0.4, 0, 0);
    g.CNOT(b, a, 0);
    g.RY(b, 0.6, 0, 0);
    g.CNOT(a, b, 0);
    for (int q : {a, b}) { g.RZ(q, 0.3, 0, 0); g.RY(q, 0.5, 0, 0); g.RZ(q, 0.7, 0, 0); }
}

static int cmd_run(int argc, char **argv)
{
    if (argc < 6) return 2;
    Circuit c;
    if (!load_circuit(argv[2], c)) { std::fprintf(stderr, "cannot read %s\n", argv[2]); return 1; }
    std::string prefix = argv[3];
    long seed = std::atol(argv[4]);
    long shots = std::atol(argv[5]);
    std::memset(&g_tm, 0, sizeof g_tm);
    g_tm.tm_year = (int)(seed - 70);

    CPUQuantumGates g;
    QuantumGateParam param;
    param.setQBitNum(c.n);
    if (g.initState(&param) != qErrorNone) return 1;

    bool dumped = false;
    std::vector<uint64_t> out;
    auto dump = [&]() {
        FILE *f = std::fopen((prefix + ".state").c_str(), "wb");
        std::fwrite(g.mvQuantumStat.data(), 16, g.mvQuantumStat.size(), f);
        std::fclose(f);
        dumped = true;
    };
    long nshots = shots > 0 ? shots : 1;
    for (long s = 0; s < nshots; s++) {
        g.endGate(&param, &g);
        std::map<size_t, int> creg;
        for (const qsv_gate &r : c.gates) {
            if (r.kind == QSV_GATE_MEASURE && !dumped) dump();
            if (!apply(g, r, creg)) { std::fprintf(stderr, "kind %d not expressible\n", r.kind); return 1; }
        }
        if (!dumped) dump();
        uint64_t bits = 0;
        for (auto &kv : creg) bits |= (uint64_t)(kv.second & 1) << kv.first;
        out.push_back(bits);
    }
    if (shots > 0) {
        FILE *f = std::fopen((prefix + ".shots").c_str(), "wb");
        std::fwrite(out.data(), 8, out.size(), f);
        std::fclose(f);
    }
    return 0;
}

static int cmd_time(int argc, char **argv)
{
    if (argc < 5) return 2;
    Circuit c;
    if (!load_circuit(argv[2], c)) { std::fprintf(stderr, "cannot read %s\n", argv[2]); return 1; }
    long stride = std::atol(argv[3]);
    double max_s = std::atof(argv[4]);
    if (stride < 1) stride = 1;
    std::memset(&g_tm, 0, sizeof g_tm);
    g_tm.tm_year = 1 - 70;

    CPUQuantumGates g;
    QuantumGateParam param;
    param.setQBitNum(c.n);
    auto t0 = std::chrono::steady_clock::now();
    if (g.initState(&param) != qErrorNone) return 1;
    auto t1 = std::chrono::steady_clock::now();
    std::map<size_t, int> creg;
    long done = 0, sweeps = 0;
    for (size_t i = 0; i < c.gates.size(); i += (size_t)stride) {
        const qsv_gate &r = c.gates[i];
        if (r.kind == QSV_GATE_U2) { apply_u2_template(g, r.q0, r.q1); sweeps += 18; }
        else if (r.kind == QSV_GATE_U1) { g.RZ(r.q0, 0.1, 0, 0); g.RY(r.q0, 0.2, 0, 0); g.RZ(r.q0, 0.3, 0, 0); sweeps += 3; }
        else { if (!apply(g, r, creg)) return 1; sweeps += 1; }
        done++;
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count() > max_s) break;
    }
    auto t2 = std::chrono::steady_clock::now();
    double init_s = std::chrono::duration<double>(t1 - t0).count();
    double run_s = std::chrono::duration<double>(t2 - t1).count();
    const char *omp = std::getenv("OMP_NUM_THREADS");
    std::printf("{\"n\": %d, \"gates\": %ld, \"stride\": %ld, \"reference_calls\": %ld, \"seconds\": %.6f, "
                "\"init_seconds\": %.6f, \"omp_num_threads\": \"%s\"}\n",
                c.n, done, stride, sweeps, run_s, init_s, omp ? omp : "");
    return 0;
}

static int cmd_check(int argc, char **argv)
{
    if (argc < 4) return 2;
    Circuit c;
    if (!load_circuit(argv[2], c)) { std::fprintf(stderr, "cannot read %s\n", argv[2]); return 1; }
    const long nrec = std::atol(argv[3]);
    std::memset(&g_tm, 0, sizeof g_tm);
    g_tm.tm_year = 1 - 70;
    CPUQuantumGates g;
    QuantumGateParam param;
    param.setQBitNum(c.n);
    if (g.initState(&param) != qErrorNone) return 1;
    std::map<size_t, int> creg;
    long done = 0;
    for (size_t i = 0; i < c.gates.size() && done < nrec; i++, done++) {
        const qsv_gate &r = c.gates[i];
        if (r.kind == QSV_GATE_U2) apply_u2_template(g, r.q0, r.q1);
        else if (r.kind == QSV_GATE_U1) { g.RZ(r.q0, 0.1, 0, 0); g.RY(r.q0, 0.2, 0, 0); g.RZ(r.q0, 0.3, 0, 0); }
        else if (!apply(g, r, creg)) return 1;
    }
    uint64_t h = 1469598103934665603ull;
    const unsigned char *b = reinterpret_cast<const unsigned char *>(g.mvQuantumStat.data());
    const size_t nb = g.mvQuantumStat.size() * 16;
    for (size_t i = 0; i < nb; i++) h = (h ^ b[i]) * 1099511628211ull;
    const char *omp = std::getenv("OMP_NUM_THREADS");
    std::printf("{\"n\": %d, \"records\": %ld, \"fnv1a\": \"%016llx\", \"omp_num_threads\": \"%s\"}\n", c.n, done,
                (unsigned long long)h, omp ? omp : "");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "check") == 0) return cmd_check(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "run") == 0) return cmd_run(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "time") == 0) return cmd_time(argc, argv);
    std::fprintf(stderr, "usage: %s run CIRCUIT OUTPREFIX SEED SHOTS | time CIRCUIT STRIDE MAX_SECONDS | "
                         "check CIRCUIT NRECORDS\n", argv[0]);
    return 2;
}
