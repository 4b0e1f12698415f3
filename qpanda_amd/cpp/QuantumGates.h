// QuantumGates.h -- the gate-API interface the engine plugs into.
//
// Same class, method names, signatures, error codes and parameter object as
// QPanda's abstract back end (QPandaSDK/QuantumInstructionHandle/QuantumGates.h:31-474,
// QError.h:3-17, QuantumGateParameter.h:20-49), declared here so the adapter and
// its tests build without the QPanda tree.  Inside QPanda the adapter compiles
// against QPanda's own header instead; the two declarations must stay identical
// for HipQuantumGates to be a drop-in back end.
#ifndef QSV_QUANTUM_GATES_H
#define QSV_QUANTUM_GATES_H

#include <complex>
#include <cstddef>
#include <map>
#include <string>
#include <utility>
#include <vector>

enum QError {
    qErrorNone = 2,
    undefineError,
    qParameterError,
    qbitError,
    loadFileError,
    initStateError,
    destroyStateError,
    setComputeUnitError,
    runProgramError,
    getResultError,
    getQStateError,
    notReadFileError
};

class QuantumGateParam {
public:
    size_t mPMeasureSize;
    int mQuantumBitNumber;
    std::map<size_t, int> mReturnValue;
    std::vector<std::pair<size_t, double>> mPMeasure;

    QuantumGateParam() : mPMeasureSize(0), mQuantumBitNumber(0) {}
    bool setQBitNum(int n)
    {
        mQuantumBitNumber = n;
        return true;
    }
};

typedef std::complex<double> COMPLEX;
typedef std::vector<size_t> Qnum;
typedef std::vector<std::complex<double>> QStat;

class QuantumGates {
public:
    QuantumGates();
    virtual ~QuantumGates() = 0;

    // Park-Miller "minimal standard" generator seeded once per process from
    // the local time (QuantumGates.cpp:46-85).
    double randGenerator();

    virtual size_t getQStateSize() = 0;
    virtual bool getQState(std::string &sState, QuantumGateParam *pQuantumProParam) = 0;
    virtual QError Hadamard(size_t qn, double error_rate) = 0;
    virtual QError Reset(size_t qn) = 0;
    virtual QError pMeasure(Qnum &qnum, std::vector<std::pair<size_t, double>> &mResult) = 0;
    virtual QError Hadamard(size_t qn, double error_rate, Qnum vControlBit, size_t stQuantumBitNumber) = 0;
    virtual QError RX(size_t qn, double theta, double error_rate, int iDagger) = 0;
    virtual QError RX(size_t qn, double theta, double error_rate, Qnum vControlBitNumber,
                      size_t stQuantumBitNumber, int iDagger) = 0;
    virtual QError RY(size_t qn, double theta, double error_rate, int iDagger) = 0;
    virtual QError RY(size_t qn, double theta, double error_rate, Qnum vControlBit,
                      size_t stQuantumBitNumber, int iDagger) = 0;
    virtual QError RZ(size_t qn, double theta, double error_rate, int iDagger) = 0;
    virtual QError RZ(size_t qn, double theta, double error_rate, Qnum vControlBitNumber,
                      size_t stQuantumBitNumber, int iDagger) = 0;
    virtual QError CNOT(size_t qn_1, size_t qn_2, double error_rate) = 0;
    virtual QError CNOT(size_t qn_1, size_t qn_2, double error_rate, Qnum vControlBitNumber,
                        int stQuantumBitNumber) = 0;
    virtual QError CR(size_t qn_1, size_t qn_2, double theta, double error_rate, int iDagger) = 0;
    virtual QError iSWAP(size_t qn_1, size_t qn_2, double error_rate, int iDagger) = 0;
    virtual QError iSWAP(size_t qn_1, size_t qn_2, double error_rate, Qnum vControlBitNumber,
                         int stQuantumBitNumber, int iDagger) = 0;
    virtual QError sqiSWAP(size_t qn_1, size_t qn_2, double error_rate, int iDagger) = 0;
    virtual QError sqiSWAP(size_t qn_1, size_t qn_2, double error_rate, Qnum vControlBitNumber,
                           int stQuantumBitNumber, int iDagger) = 0;
    virtual QError controlSwap(size_t qn_1, size_t qn_2, size_t qn_3, double error_rate) = 0;
    virtual int qubitMeasure(size_t qn) = 0;
    virtual QError unitarySingleQubitGate(size_t qn, void *matrix, double error_rate) = 0;
    virtual QError initState(QuantumGateParam *pQuantumProParam) = 0;
    virtual bool compareCalculationUnitType(std::string &sCalculationUnitType) = 0;
    virtual QError NOT(size_t qn, double error_rate) = 0;
    virtual QError NOT(size_t qn, double error_rate, Qnum vControlBit, int stQuantumBitNumber) = 0;
    virtual QError toffoli(size_t stControlBit1, size_t stControlBit2, size_t stQuantumBit,
                           double errorRate, int stQuantumBitNumber) = 0;
    virtual QError destroyState(size_t) { return qErrorNone; }
    virtual QError endGate(QuantumGateParam *pQuantumProParam, QuantumGates *pQGate) = 0;

protected:
    virtual QError matReverse(COMPLEX (*)[4], COMPLEX (*)[4]) = 0;
};

#endif
