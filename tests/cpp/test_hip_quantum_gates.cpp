// test_hip_quantum_gates.cpp -- the known-answer tests of QPanda's CPU back end
// (QPandaSDK/testMain/testCPUQuantumGates.cpp, tolerance ZERO = 1e-12, :6),
// run against HipQuantumGates on the GPU through the QuantumGates interface.
// Each test names the reference test it follows.  Exit status = failures.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "HipQuantumGates.h"

#define PI 3.141592653589793
#define ZERO 1e-12
#define SQ2 0.707106781186548
#define REPEAT 20

static int g_fail = 0, g_checks = 0;
#define EXPECT(cond)                                                              \
    do {                                                                          \
        g_checks++;                                                               \
        if (!(cond)) {                                                            \
            g_fail++;                                                             \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
        }                                                                         \
    } while (0)

static void init(HipQuantumGates &g, int n)
{
    QuantumGateParam p;
    p.setQBitNum(n);
    EXPECT(g.initState(&p) == qErrorNone);
}

// three distinct qubits in [0, 10) from the back end's own generator
static void triple(HipQuantumGates &g, size_t &a, size_t &b, size_t &c)
{
    a = (size_t)(g.randGenerator() * 10);
    b = (size_t)(g.randGenerator() * 10);
    c = (size_t)(g.randGenerator() * 10);
    while (a == b) b = (size_t)(g.randGenerator() * 10);
    while (a == c || b == c) c = (size_t)(g.randGenerator() * 10);
}

static void testInitState()  // :14-31
{
    HipQuantumGates g;
    init(g, 10);
    QStat s = g.state();
    double sum = 0;
    for (auto &a : s) sum += std::norm(a);
    EXPECT(std::abs(sum - 1) < ZERO && std::abs(s[0].real() - 1) < ZERO);
    EXPECT(g.getQStateSize() == 1024u);
}

static void testReset()  // :33-48
{
    HipQuantumGates g;
    init(g, 10);
    for (int i = 0; i < 10; i++) {
        EXPECT(g.RX(i, g.randGenerator(), 0, 0) == qErrorNone);
        EXPECT(g.Reset(i) == qErrorNone);
    }
    EXPECT(std::abs(std::abs(g.state()[0]) - 1) < ZERO);
}

static void testNOT()  // :50-64
{
    HipQuantumGates g;
    init(g, 10);
    for (int i = 0; i < 10; i++) {
        EXPECT(g.NOT(i, 0) == qErrorNone);
        EXPECT(std::abs(g.state()[1ll << i].real() - 1) < ZERO);
        EXPECT(g.NOT(i, 0) == qErrorNone);
        EXPECT(std::abs(g.state()[0].real() - 1) < ZERO);
    }
}

static void testHadamard()  // :66-91
{
    HipQuantumGates g;
    init(g, 10);
    for (int i = 0; i < 10; i++) {
        g.Hadamard(i, 0);
        QStat s = g.state();
        EXPECT(std::abs(s[0].real() - SQ2) < ZERO && std::abs(s[1ll << i].real() - SQ2) < ZERO);
        g.Hadamard(i, 0);
    }
    for (int i = 0; i < 10; i++) {
        g.NOT(i, 0);
        g.Hadamard(i, 0);
        QStat s = g.state();
        EXPECT(std::abs(s[0].real() - SQ2) < ZERO && std::abs(s[1ll << i].real() + SQ2) < ZERO);
        g.Hadamard(i, 0);
        g.NOT(i, 0);
    }
}

static void testRotations()  // RX :93-118, RY :120-145, RZ :147-176
{
    HipQuantumGates g;
    init(g, 10);
    const double t = g.randGenerator() * 2 * PI;
    for (int i = 0; i < 10; i++) {
        g.RX(i, t, 0, 0);
        QStat s = g.state();
        EXPECT(std::abs(s[0].real() - std::cos(t / 2)) < ZERO);
        EXPECT(std::abs(s[1ll << i].imag() + std::sin(t / 2)) < ZERO);
        g.RX(i, t, 0, 1);
        g.RY(i, t, 0, 0);
        s = g.state();
        EXPECT(std::abs(s[0].real() - std::cos(t / 2)) < ZERO);
        EXPECT(std::abs(s[1ll << i].real() - std::sin(t / 2)) < ZERO);
        g.RY(i, t, 0, 1);
        g.Hadamard(i, 0);
        g.RZ(i, t, 0, 0);
        s = g.state();
        EXPECT(std::abs(s[0] - SQ2 * std::exp(COMPLEX(0, -t / 2))) < ZERO);
        EXPECT(std::abs(s[1ll << i] - SQ2 * std::exp(COMPLEX(0, t / 2))) < ZERO);
        g.RZ(i, t, 0, 1);
        g.Hadamard(i, 0);
    }
    EXPECT(std::abs(g.state()[0] - COMPLEX(1, 0)) < 1e-12);
}

static void testCNOT()  // :178-210
{
    HipQuantumGates g;
    init(g, 10);
    for (int i = 0; i < 10; i++)
        for (int j = 0; j < 10; j++) {
            if (i == j) continue;
            g.CNOT(i, j, 0);
            EXPECT(std::abs(g.state()[0].real() - 1) < ZERO);
            g.NOT(j, 0);
            g.CNOT(i, j, 0);
            EXPECT(std::abs(g.state()[1ll << j].real() - 1) < ZERO);
            g.NOT(i, 0);
            g.CNOT(i, j, 0);
            EXPECT(std::abs(g.state()[1ll << i].real() - 1) < ZERO);
            g.CNOT(i, j, 0);
            EXPECT(std::abs(g.state()[(1ll << i) + (1ll << j)].real() - 1) < ZERO);
            g.NOT(i, 0);
            g.NOT(j, 0);
        }
}

static void testCR()  // :212-248
{
    HipQuantumGates g;
    init(g, 10);
    const double t = g.randGenerator();
    for (int i = 0; i < 10; i++)
        for (int j = 0; j < 10; j++) {
            if (i == j) continue;
            g.NOT(i, 0);
            g.NOT(j, 0);
            g.CR(i, j, t, 0, 0);
            const COMPLEX v = g.state()[(1ll << i) + (1ll << j)];
            EXPECT(std::abs(v.real() - std::cos(t)) < ZERO && std::abs(v.imag() - std::sin(t)) < ZERO);
            g.CR(i, j, t, 0, 1);
            g.NOT(i, 0);
            g.NOT(j, 0);
        }
}

static void testISWAP()  // :250-333
{
    HipQuantumGates g;
    init(g, 10);
    for (int i = 0; i < 10; i++)
        for (int j = 0; j < 10; j++) {
            if (i == j) continue;
            g.NOT(j, 0);
            g.iSWAP(i, j, 0, 0);
            EXPECT(std::abs(g.state()[1ll << i].imag() - 1) < ZERO);
            g.iSWAP(i, j, 0, 1);
            g.sqiSWAP(i, j, 0, 0);
            QStat s = g.state();
            EXPECT(std::abs(s[1ll << i].imag() - SQ2) < ZERO && std::abs(s[1ll << j].real() - SQ2) < ZERO);
            g.sqiSWAP(i, j, 0, 1);
            g.NOT(j, 0);
            EXPECT(std::abs(g.state()[0].real() - 1) < 1e-12);
        }
}

static void testQubitMeasure()  // :335-358
{
    HipQuantumGates g;
    init(g, 4);
    int n0 = 0;
    const int times = 1 << 16;
    for (int i = 0; i < times; i++) {
        g.Hadamard(0, 0);
        if (!g.qubitMeasure(0)) n0++;
    }
    EXPECT(std::abs(n0 * 1.0 / times - 0.5) < 1e-2);
}

static void testPMeasure()  // :360-378
{
    HipQuantumGates g;
    init(g, 10);
    for (int i = 0; i < 10; i++) g.Hadamard(i, 0);
    Qnum q = {2, 1, 3};
    std::vector<std::pair<size_t, double>> out;
    EXPECT(g.pMeasure(q, out) == qErrorNone);
    EXPECT(out.size() == 8u);
    for (auto &p : out) EXPECT(std::abs(p.second - 1.0 / 8) < ZERO);
}

static void testToffoli()  // :380-442
{
    HipQuantumGates g;
    init(g, 4);
    for (int c1 = 0; c1 < 4; c1++)
        for (int c2 = 0; c2 < 4; c2++)
            for (int t = 0; t < 4; t++) {
                if (c1 == c2 || c1 == t || c2 == t) continue;
                for (int v = 0; v < 8; v++) {
                    g.endGate(nullptr, &g);
                    const int b0 = v & 1, b1 = v >> 1 & 1, b2 = v >> 2 & 1;
                    if (b0) g.NOT(c1, 0);
                    if (b1) g.NOT(c2, 0);
                    if (b2) g.NOT(t, 0);
                    g.toffoli(c1, c2, t, 0, 4);
                    const int bt = b2 ^ (b0 & b1);
                    EXPECT(std::abs(g.state()[(b0 << c1) | (b1 << c2) | (bt << t)].real() - 1) < ZERO);
                }
            }
}

static void testControlled()  // :448-731 (controls, then the target last)
{
    HipQuantumGates g;
    init(g, 10);
    for (int rep = 0; rep < REPEAT; rep++) {
        size_t a, b, c;
        triple(g, a, b, c);
        Qnum list = {a, b, c};
        const long long i11 = (1ll << a) + (1ll << b), i111 = i11 + (1ll << c);
        const double t = g.randGenerator() * 2 * PI;
        // NOT: |110> <-> |111>
        g.NOT(c, 0, list, 10);
        EXPECT(std::abs(g.state()[0].real() - 1) < ZERO);
        g.NOT(a, 0);
        g.NOT(b, 0);
        g.NOT(c, 0, list, 10);
        EXPECT(std::abs(g.state()[i111].real() - 1) < ZERO);
        g.NOT(c, 0, list, 10);
        // Hadamard: |110> -> SQ2 |110> + SQ2 |111>
        g.Hadamard(c, 0, list, 10);
        QStat s = g.state();
        EXPECT(std::abs(s[i11].real() - SQ2) < ZERO && std::abs(s[i111].real() - SQ2) < ZERO);
        g.Hadamard(c, 0, list, 10);
        // RX: cos, -i sin
        g.RX(c, t, 0, list, 10, 0);
        s = g.state();
        EXPECT(std::abs(s[i11].real() - std::cos(t / 2)) < ZERO && std::abs(s[i111].imag() + std::sin(t / 2)) < ZERO);
        g.RX(c, t, 0, list, 10, 1);
        // RY: cos, sin
        g.RY(c, t, 0, list, 10, 0);
        s = g.state();
        EXPECT(std::abs(s[i11].real() - std::cos(t / 2)) < ZERO && std::abs(s[i111].real() - std::sin(t / 2)) < ZERO);
        g.RY(c, t, 0, list, 10, 1);
        // RZ on |11>(|0> + |1>)/sqrt2: e^{-it/2}, e^{it/2}
        g.Hadamard(c, 0, list, 10);
        g.RZ(c, t, 0, list, 10, 0);
        s = g.state();
        EXPECT(std::abs(s[i11] - SQ2 * std::exp(COMPLEX(0, -t / 2))) < ZERO);
        EXPECT(std::abs(s[i111] - SQ2 * std::exp(COMPLEX(0, t / 2))) < ZERO);
        g.RZ(c, t, 0, list, 10, 1);
        g.Hadamard(c, 0, list, 10);
        // CNOT with a control list forwards to the controlled NOT
        g.CNOT(b, c, 0, list, 10);
        EXPECT(std::abs(g.state()[i111].real() - 1) < ZERO);
        g.CNOT(b, c, 0, list, 10);
        g.NOT(a, 0);
        g.NOT(b, 0);
        EXPECT(std::abs(g.state()[0].real() - 1) < 1e-12);
    }
}

static void testShorAlgorithm()  // :733-772
{
    HipQuantumGates g;
    init(g, 12);
    g.NOT(0, 0);
    for (int i = 4; i < 12; i++) g.Hadamard(i, 0);
    for (int i = 0; i < 4; i++) g.CNOT(i, 4, 0);
    g.controlSwap(4, 2, 1, 0);
    g.controlSwap(4, 2, 3, 0);
    g.controlSwap(4, 3, 0, 0);
    g.controlSwap(5, 2, 0, 0);
    g.controlSwap(5, 3, 1, 0);
    for (int i = 11; i >= 4; i--) {
        g.Hadamard(i, 0);
        for (int j = i - 1; j >= 4; j--) g.CR(i, j, -PI / (1ll << (i - j)), 0, 0);
    }
    Qnum q = {5, 4};
    std::vector<std::pair<size_t, double>> out;
    EXPECT(g.pMeasure(q, out) == qErrorNone);
    EXPECT(out.size() == 4u);
    for (auto &p : out) EXPECT(std::abs(p.second - 0.25) < ZERO);
}

static void testSeededDrawsAndText()
{
    // seeded stream: the first draw equals the Park-Miller successor of the seed
    HipQuantumGates g;
    init(g, 3);
    g.seedRandom(1);
    g.Hadamard(0, 0.0);                 // consumes draw 1 = 16807 / (2^31 - 1)
    g.Hadamard(1, 16807.0 / 2147483647 + 1e-9);  // draw 2 = 282475249/.. > err: applied
    QStat s = g.state();
    EXPECT(std::abs(s[3].real() - SQ2 * SQ2) < ZERO);
    std::string txt;
    EXPECT(g.getQState(txt, nullptr));
    EXPECT(txt.find("state[0].real = ") == 0);
    // error_rate above the draw skips the gate (CPUQuantumGates.cpp:97)
    g.seedRandom(1);
    g.NOT(2, 0.5);                      // draw 7.8e-6 <= 0.5: skipped
    EXPECT(std::abs(g.state()[3].real() - SQ2 * SQ2) < ZERO);
    std::string unit = "HIP";
    EXPECT(g.compareCalculationUnitType(unit));
}

int main()
{
    const std::vector<std::pair<const char *, std::function<void()>>> tests = {
        {"testInitState", testInitState}, {"testReset", testReset}, {"testNOT", testNOT},
        {"testHadamard", testHadamard}, {"testRotations", testRotations}, {"testCNOT", testCNOT},
        {"testCR", testCR}, {"testISWAP", testISWAP}, {"testQubitMeasure", testQubitMeasure},
        {"testPMeasure", testPMeasure}, {"testToffoli", testToffoli}, {"testControlled", testControlled},
        {"testShorAlgorithm", testShorAlgorithm}, {"testSeededDrawsAndText", testSeededDrawsAndText},
    };
    for (auto &t : tests) {
        const int before = g_fail;
        std::printf("[ RUN  ] %s\n", t.first);
        std::fflush(stdout);
        t.second();
        std::printf("[ %s ] %s\n", g_fail == before ? " OK " : "FAIL", t.first);
        std::fflush(stdout);
    }
    std::printf("%d checks, %d failures\n", g_checks, g_fail);
    return g_fail;
}
