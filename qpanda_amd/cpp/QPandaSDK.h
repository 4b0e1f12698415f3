// QPandaSDK.h -- the QRunes program runner of QPanda (QPandaSDK/QPandaAPI/QPandaAPI.h:35-120)
// on the MI355X engine.
//
// Same methods, return codes and result strings as the reference's QPandaSDK:
//   loadFile     parse a QRunes file (QPandaAPI.cpp:58-95): qParameterError for
//                an empty path, qbitError when parsing fails or the program
//                declares fewer than 1 or more than 30 qubits;
//   setComputeUnit  2 (GPU) selects this engine; 1 (CPU) is the reference's
//                CPU simulator, which this library does not contain
//                (setComputeUnitError); other values qParameterError;
//   run(n)       n > 0 runs of the program from |0...0> (QPandaAPI.cpp:123-205);
//                MEASURE programs count the classical-register strings
//                (creg 0 rightmost) and report frequencies; a PMEASURE program
//                reports the top-10 probabilities under MSB-first bit strings;
//   getResult / getQuantumState  the result lines "<key> <value>\n" and the
//                state text of the last run.
// Every gate call draws once from the Park-Miller stream, as in the reference.
//
// Runs whose measurements all come after the last gate (no RESET, distinct
// measured qubits) do not re-run the program per shot: the gates run once and
// the shots are sampled from that state (qsv_sample_shots) with the same draws
// the re-runs would consume; the last shot's measurements are then replayed
// on the state so the final state and registers are the last run's.  The
// outcomes are identical to re-running (tests compare both ways).
#ifndef QSV_QPANDA_SDK_H
#define QSV_QPANDA_SDK_H

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "HipQuantumGates.h"
#include "QRunes.h"

namespace QPanda {

class QPandaSDK {
public:
    std::string msFileName;

    QPandaSDK();
    ~QPandaSDK();
    QPandaSDK(const QPandaSDK &) = delete;
    QPandaSDK &operator=(const QPandaSDK &) = delete;

    static QPandaSDK *getIntance();

    std::string getResult();
    std::string getQuantumState();
    int loadFile(const std::string &sFilePath);
    int setComputeUnit(int iComputeUnit);
    int run(int iRepeat);

    // Not part of the reference class.
    // Seed the Park-Miller stream of the compute unit (default: the
    // reference's time-based seed expression, at setComputeUnit).
    void setRandomSeed(int32_t seed);
    // 1 when the last run sampled its shots from one execution, 0 when it
    // re-ran the program per shot.
    int lastRunSampled() const { return mSampled ? 1 : 0; }
    // Force re-running per shot (for comparisons).
    void setResampleDisabled(bool off) { mNoSampling = off; }
    const qsvq::QRProgram &program() const { return mProg; }
    // FP64 amplitudes of the last run's final state (empty before a run).
    QStat amplitudes() const;
    // The reference CLI's state files (QCommand.cpp:402-487): path + ".fullstate"
    // (the getQuantumState text) or path + ".fullstate.binary" ({f64 re, f64 im}
    // records), streamed from the device.  qErrorNone, or qParameterError for
    // an empty path, runProgramError before a run or when writing fails.
    int saveState(const std::string &path, bool binary);

private:
    bool transmit();
    bool transmitGatesOnly(size_t count);
    bool execute(const qsvq::QRNode &n);
    bool samplingApplies(size_t *first_measure) const;
    int runSampled(int iRepeat, size_t first_measure);
    void countState();
    void countWords(const std::vector<uint64_t> &words, size_t first_measure);
    static void integerToBinary(size_t number, std::string &out, size_t len);

    std::map<std::string, double> mQResultMap;
    qsvq::QRProgram mProg;
    bool mbIsRead = false;
    bool mbIsRun = false;
    int miMeasureType = -1;
    QuantumGateParam *mQGatesParam = nullptr;
    HipQuantumGates *mGates = nullptr;
    bool mStateReady = false;
    bool mSampled = false;
    bool mNoSampling = false;
};

}  // namespace QPanda

#endif
