// HipQuantumGates.h -- QPanda back end on the MI355X engine.
//
// Implements the QuantumGates interface (QPandaSDK/QuantumInstructionHandle/QuantumGates.h)
// over the C ABI of include/qsv.h, with the CPU simulator's semantics
// (CPUQuantumGates.cpp): one random draw per gate call, applied when the draw
// exceeds error_rate; control lists hold the controls followed by the target
// (Instruction.cpp:62); RZ = diag(e^{-i t/2}, e^{i t/2}); iSWAP phase +i;
// measurement compares (float)draw with P(0) and renormalises by 1/sqrt(p).
// A drop-in for QuantumInstructionHandleAPI::setComputeUnit
// (QuantumInstructionHandleAPI.cpp:194-229) next to CPUQuantumGates.
#ifndef QSV_HIP_QUANTUM_GATES_H
#define QSV_HIP_QUANTUM_GATES_H

#include <cstdint>

#include "QuantumGates.h"

struct qsv_state;

class HipQuantumGates : public QuantumGates {
public:
    HipQuantumGates();
    ~HipQuantumGates();

    size_t getQStateSize();
    bool getQState(std::string &sState, QuantumGateParam *pQuantumProParam);
    QError Hadamard(size_t qn, double error_rate);
    QError Reset(size_t qn);
    QError pMeasure(Qnum &qnum, std::vector<std::pair<size_t, double>> &mResult);
    QError Hadamard(size_t qn, double error_rate, Qnum vControlBit, size_t stQuantumBitNumber);
    QError RX(size_t qn, double theta, double error_rate, int iDagger);
    QError RX(size_t qn, double theta, double error_rate, Qnum vControlBitNumber, size_t stQuantumBitNumber,
              int iDagger);
    QError RY(size_t qn, double theta, double error_rate, int iDagger);
    QError RY(size_t qn, double theta, double error_rate, Qnum vControlBit, size_t stQuantumBitNumber,
              int iDagger);
    QError RZ(size_t qn, double theta, double error_rate, int iDagger);
    QError RZ(size_t qn, double theta, double error_rate, Qnum vControlBitNumber, size_t stQuantumBitNumber,
              int iDagger);
    QError CNOT(size_t qn_1, size_t qn_2, double error_rate);
    QError CNOT(size_t qn_1, size_t qn_2, double error_rate, Qnum vControlBitNumber, int stQuantumBitNumber);
    QError CR(size_t qn_1, size_t qn_2, double theta, double error_rate, int iDagger);
    QError iSWAP(size_t qn_1, size_t qn_2, double error_rate, int iDagger);
    QError iSWAP(size_t qn_1, size_t qn_2, double error_rate, Qnum vControlBitNumber, int stQuantumBitNumber,
                 int iDagger);
    QError sqiSWAP(size_t qn_1, size_t qn_2, double error_rate, int iDagger);
    QError sqiSWAP(size_t qn_1, size_t qn_2, double error_rate, Qnum vControlBitNumber, int stQuantumBitNumber,
                   int iDagger);
    QError controlSwap(size_t qn_1, size_t qn_2, size_t qn_3, double error_rate);
    int qubitMeasure(size_t qn);
    QError unitarySingleQubitGate(size_t qn, void *matrix, double error_rate);
    QError initState(QuantumGateParam *pQuantumProParam);
    bool compareCalculationUnitType(std::string &sCalculationUnitType);
    QError NOT(size_t qn, double error_rate);
    QError NOT(size_t qn, double error_rate, Qnum vControlBit, int stQuantumBitNumber);
    QError toffoli(size_t stControlBit1, size_t stControlBit2, size_t stQuantumBit, double errorRate,
                   int stQuantumBitNumber);
    QError destroyState(size_t stQNum);
    QError endGate(QuantumGateParam *pQuantumProParam, QuantumGates *pQGate);

    // Not part of the reference interface.
    // The state as a host vector (the CPU back end exposes mvQuantumStat).
    QStat state();
    // Use a private Park-Miller stream starting at `seed` instead of the
    // process-wide time-seeded one (for reproducible runs).
    void seedRandom(int32_t seed);
    // State of that private stream and the number of draws taken from it
    // (both valid after seedRandom).
    int32_t rngState() const { return mRng; }
    void setRngState(int32_t s) { mRng = s; }
    int64_t draws() const { return mDraws; }
    qsv_state *handle() { return mState; }

protected:
    QError matReverse(COMPLEX (*mat)[4], COMPLEX (*mat_rev)[4]);

private:
    double draw();
    QError status(int rc);
    // mask of the list entries other than the target
    uint64_t controlMask(const Qnum &list, size_t target, bool *ok);

    qsv_state *mState = nullptr;
    int mQubits = 0;
    bool mSeeded = false;
    int32_t mRng = 1;
    int64_t mDraws = 0;
    std::string sCalculationUnitType = "HIP";
};

#endif
