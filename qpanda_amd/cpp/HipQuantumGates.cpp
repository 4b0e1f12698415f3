// HipQuantumGates.cpp -- QuantumGates on the HIP engine (see HipQuantumGates.h).
#include "HipQuantumGates.h"

#include <algorithm>
#include <cmath>
#include <sstream>

#include "../../include/qsv.h"

HipQuantumGates::HipQuantumGates() {}

HipQuantumGates::~HipQuantumGates()
{
    if (mState) qsv_destroy(mState);
}

double HipQuantumGates::draw()
{
    if (!mSeeded) return randGenerator();
    mRng = qsv_pm_next(mRng);
    mDraws++;
    return (double)mRng / 2147483647;
}

void HipQuantumGates::seedRandom(int32_t seed)
{
    mSeeded = true;
    mRng = seed;
}

QError HipQuantumGates::status(int rc)
{
    if (rc == QSV_OK) return qErrorNone;
    if (rc == QSV_ERR_QUBIT) return qbitError;
    if (rc == QSV_ERR_ARG) return qParameterError;
    return undefineError;
}

uint64_t HipQuantumGates::controlMask(const Qnum &list, size_t target, bool *ok)
{
    uint64_t m = 0;
    *ok = true;
    for (size_t q : list) {
        if (q >= (size_t)mQubits) *ok = false;
        else if (q != target) m |= uint64_t(1) << q;
    }
    return m;
}

size_t HipQuantumGates::getQStateSize() { return mState ? (size_t(1) << mQubits) : 0; }

QStat HipQuantumGates::state()
{
    QStat out;
    if (!mState) return out;
    out.resize(size_t(1) << mQubits);
    qsv_get_state(mState, reinterpret_cast<double *>(out.data()));
    return out;
}

// Text form of CPUQuantumGates::getQState (CPUQuantumGates.cpp:64-82).
bool HipQuantumGates::getQState(std::string &sState, QuantumGateParam *)
{
    if (!mState) return false;
    QStat st = state();
    std::stringstream ss;
    for (size_t i = 0; i < st.size(); i++)
        ss << "state[" << i << "].real = " << st[i].real() << " "
           << "state[" << i << "].imag = " << st[i].imag() << "\n";
    sState.append(ss.str());
    return true;
}

QError HipQuantumGates::initState(QuantumGateParam *p)
{
    if (!p) return undefineError;
    const int n = p->mQuantumBitNumber;
    if (mState && mQubits == n) return status(qsv_reset_zero(mState));
    if (mState) {
        qsv_destroy(mState);
        mState = nullptr;
    }
    if (qsv_create(n, QSV_C128, &mState) != QSV_OK) {
        mState = nullptr;
        return undefineError;
    }
    // Runs of gates that share a tile go through HBM once; the next readout
    // (measure, pMeasure, getQState, ...) applies what is queued, and the
    // state is bit-identical to one sweep per gate.
    if (qsv_set_gate_tiling(mState, 1) != QSV_OK) {
        qsv_destroy(mState);
        mState = nullptr;
        return undefineError;
    }
    mQubits = n;
    return qErrorNone;
}

QError HipQuantumGates::destroyState(size_t)
{
    if (mState) qsv_destroy(mState);
    mState = nullptr;
    mQubits = 0;
    return qErrorNone;
}

QError HipQuantumGates::endGate(QuantumGateParam *, QuantumGates *)
{
    if (!mState) return undefineError;
    return status(qsv_reset_zero(mState));
}

bool HipQuantumGates::compareCalculationUnitType(std::string &s) { return s == sCalculationUnitType; }

// ---- single-qubit gates: draw, then apply when the draw exceeds error_rate
QError HipQuantumGates::Hadamard(size_t qn, double err)
{
    if (draw() > err) return status(qsv_h(mState, (int)qn, 0, 0));
    return qErrorNone;
}

QError HipQuantumGates::NOT(size_t qn, double err)
{
    if (draw() > err) return status(qsv_x(mState, (int)qn, 0));
    return qErrorNone;
}

QError HipQuantumGates::RX(size_t qn, double theta, double err, int dag)
{
    if (draw() > err) return status(qsv_rx(mState, (int)qn, theta, dag, 0));
    return qErrorNone;
}

QError HipQuantumGates::RY(size_t qn, double theta, double err, int dag)
{
    if (draw() > err) return status(qsv_ry(mState, (int)qn, theta, dag, 0));
    return qErrorNone;
}

QError HipQuantumGates::RZ(size_t qn, double theta, double err, int dag)
{
    if (draw() > err) return status(qsv_rz(mState, (int)qn, theta, dag, 0));
    return qErrorNone;
}

// ---- controlled forms: the list holds the controls, then the target.  H, RX
// and RZ take the target from qn (CPUQuantumGates.cpp:202-258, :323-383,
// :584-655); RY and NOT from the list's last entry (:407, :1323).
QError HipQuantumGates::Hadamard(size_t qn, double err, Qnum ctrl, size_t)
{
    if (!(draw() > err)) return qErrorNone;
    bool ok;
    const uint64_t m = controlMask(ctrl, qn, &ok);
    if (!ok) return qbitError;
    return status(qsv_h(mState, (int)qn, m, 0));
}

QError HipQuantumGates::RX(size_t qn, double theta, double err, Qnum ctrl, size_t, int dag)
{
    if (!(draw() > err)) return qErrorNone;
    bool ok;
    const uint64_t m = controlMask(ctrl, qn, &ok);
    if (!ok) return qbitError;
    return status(qsv_rx(mState, (int)qn, theta, dag, m));
}

QError HipQuantumGates::RY(size_t, double theta, double err, Qnum ctrl, size_t, int dag)
{
    if (!(draw() > err)) return qErrorNone;
    if (ctrl.empty()) return qParameterError;
    const size_t t = ctrl.back();
    bool ok;
    const uint64_t m = controlMask(ctrl, t, &ok);
    if (!ok) return qbitError;
    return status(qsv_ry(mState, (int)t, theta, dag, m));
}

QError HipQuantumGates::RZ(size_t qn, double theta, double err, Qnum ctrl, size_t, int dag)
{
    if (!(draw() > err)) return qErrorNone;
    bool ok;
    const uint64_t m = controlMask(ctrl, qn, &ok);
    if (!ok) return qbitError;
    return status(qsv_rz(mState, (int)qn, theta, dag, m));
}

QError HipQuantumGates::NOT(size_t, double err, Qnum ctrl, int)
{
    if (!(draw() > err)) return qErrorNone;
    if (ctrl.empty()) return qParameterError;
    const size_t t = ctrl.back();
    bool ok;
    const uint64_t m = controlMask(ctrl, t, &ok);
    if (!ok) return qbitError;
    return status(qsv_x(mState, (int)t, m));
}

// ---- two- and three-qubit gates
QError HipQuantumGates::CNOT(size_t c, size_t t, double err)
{
    if (!(draw() > err)) return qErrorNone;
    if (c == t) return qErrorNone;                       // empty loop in the reference
    if (c >= (size_t)mQubits) return qbitError;
    return status(qsv_x(mState, (int)t, uint64_t(1) << c));
}

// CPUQuantumGates.cpp:750-757: forwards to the controlled NOT.
QError HipQuantumGates::CNOT(size_t, size_t t, double err, Qnum ctrl, int n)
{
    return NOT(t, err, ctrl, n);
}

QError HipQuantumGates::CR(size_t q1, size_t q2, double theta, double err, int dag)
{
    if (draw() > err) return status(qsv_cr(mState, (int)q1, (int)q2, theta, dag));
    return qErrorNone;
}

QError HipQuantumGates::iSWAP(size_t q1, size_t q2, double err, int dag)
{
    if (draw() > err) return status(qsv_iswap(mState, (int)q1, (int)q2, dag));
    return qErrorNone;
}

QError HipQuantumGates::sqiSWAP(size_t q1, size_t q2, double err, int dag)
{
    if (draw() > err) return status(qsv_sqiswap(mState, (int)q1, (int)q2, dag, 0));
    return qErrorNone;
}

// The CPU simulator's controlled iSWAP / sqiSWAP return without drawing or
// acting (CPUQuantumGates.cpp:914-922, :1054-1062); kept as is.
QError HipQuantumGates::iSWAP(size_t, size_t, double, Qnum, int, int) { return qErrorNone; }
QError HipQuantumGates::sqiSWAP(size_t, size_t, double, Qnum, int, int) { return qErrorNone; }

QError HipQuantumGates::controlSwap(size_t c, size_t a, size_t b, double err)
{
    if (draw() > err) return status(qsv_cswap(mState, (int)c, (int)a, (int)b));
    return qErrorNone;
}

QError HipQuantumGates::toffoli(size_t c1, size_t c2, size_t t, double err, int)
{
    if (!(draw() > err)) return qErrorNone;
    if (c1 >= (size_t)mQubits || c2 >= (size_t)mQubits) return qbitError;
    return status(qsv_x(mState, (int)t, (uint64_t(1) << c1) | (uint64_t(1) << c2)));
}

// ---- readout
int HipQuantumGates::qubitMeasure(size_t qn)
{
    int out = 0;
    if (qsv_measure(mState, (int)qn, draw(), &out) != QSV_OK) return -1;
    return out;
}

QError HipQuantumGates::Reset(size_t qn)
{
    const int o = qubitMeasure(qn);
    if (o < 0) return undefineError;
    if (o) return NOT(qn, 0);
    return qErrorNone;
}

QError HipQuantumGates::pMeasure(Qnum &qnum, std::vector<std::pair<size_t, double>> &result)
{
    std::vector<int> qs(qnum.begin(), qnum.end());
    uint64_t idx[10];
    double prob[10];
    int cnt = 0;
    const int rc = qsv_pmeasure(mState, qs.data(), (int)qs.size(), idx, prob, &cnt);
    if (rc != QSV_OK) return status(rc);
    result.clear();
    for (int i = 0; i < cnt; i++) result.emplace_back((size_t)idx[i], prob[i]);
    return qErrorNone;
}

// A 2x2 matrix given as a QStat of 4 entries, rejected unless |det| = 1
// (CPUQuantumGates.cpp:1170-1212).  Applied to qubit qn; the CPU simulator
// addresses qn - 1 (:1188), which this does not reproduce.
QError HipQuantumGates::unitarySingleQubitGate(size_t qn, void *matrix, double err)
{
    const QStat &m = *static_cast<QStat *>(matrix);
    if (m.size() < 4) return qParameterError;
    if (!(draw() > err)) return qErrorNone;
    const COMPLEX det = m[0] * m[3] - m[1] * m[2];
    if (std::abs(std::abs(det) - 1) > 0.000001) return undefineError;
    const double u[8] = {m[0].real(), m[0].imag(), m[1].real(), m[1].imag(),
                         m[2].real(), m[2].imag(), m[3].real(), m[3].imag()};
    return status(qsv_apply_1q(mState, (int)qn, u, 0));
}

// Swap the 2nd and 3rd rows and columns of a 4x4 matrix
// (CPUQuantumGates.cpp:1221-1244).
QError HipQuantumGates::matReverse(COMPLEX (*mat)[4], COMPLEX (*mat_rev)[4])
{
    static const int p[4] = {0, 2, 1, 3};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) mat_rev[i][j] = mat[p[i]][p[j]];
    return qErrorNone;
}
