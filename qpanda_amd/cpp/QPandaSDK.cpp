// QPandaSDK.cpp -- QRunes program runner on the MI355X engine.
// Run loop and result strings: QPandaAPI.cpp:58-306; executor dispatch per
// instruction: Instruction.cpp:40-827 (QuantumInstructionGet.cpp:39-124 maps
// node kinds to instructions; QuantumInstructionHandleAPI.cpp:96-118 resets
// the state before each run and stops at the first failing instruction).
#include "QPandaSDK.h"

#include <algorithm>
#include <cstring>
#include <ctime>
#include <sstream>
#include <unordered_map>
#include <vector>

#include "qsv.h"
#include "qsv_sdk.h"

namespace QPanda {

namespace {
constexpr int kCpuUnit = 1;
constexpr int kMeasureType = 1, kPMeasureType = 2;
constexpr int kMaxProgramQubits = 30;            // QPandaAPI.cpp:87
}  // namespace

using qsvq::QRNode;

QPandaSDK::QPandaSDK() = default;

QPandaSDK::~QPandaSDK()
{
    if (mGates) {
        if (mQGatesParam) mGates->destroyState(mQGatesParam->mQuantumBitNumber);
        delete mGates;
    }
    delete mQGatesParam;
}

QPandaSDK *QPandaSDK::getIntance()
{
    static QPandaSDK *instance = new QPandaSDK();
    return instance;
}

int QPandaSDK::loadFile(const std::string &sFilePath)
{
    if (sFilePath.empty()) return qParameterError;
    mbIsRead = false;
    miMeasureType = -1;
    mbIsRun = false;
    delete mQGatesParam;
    mQGatesParam = new QuantumGateParam();
    msFileName = sFilePath;
    mProg = qsvq::QRProgram();
    const int qubits = qsvq::parseQRunesFile(msFileName, mProg);
    if (qubits <= 0 || qubits > kMaxProgramQubits) return qbitError;
    mQGatesParam->mQuantumBitNumber = qubits;
    mbIsRead = true;
    return qErrorNone;
}

int QPandaSDK::setComputeUnit(int iComputeUnit)
{
    if (iComputeUnit <= 0 || iComputeUnit > 2) return qParameterError;
    if (iComputeUnit == kCpuUnit) return setComputeUnitError;   // the CPU simulator is not part of this library
    if (mGates) return qErrorNone;                              // QuantumInstructionHandleAPI.cpp:161-165
    int devices = 0;
    if (qsv_device_count(&devices) != QSV_OK || devices < 1) return setComputeUnitError;
    mGates = new HipQuantumGates();
    // the reference's time-based seed expression (QuantumGates.cpp:57-72)
    time_t raw;
    time(&raw);
    const struct tm *t = localtime(&raw);
    mGates->seedRandom(qsv_pm_seed_from_tm(t->tm_year, t->tm_mon, t->tm_mday, t->tm_hour, t->tm_min, t->tm_sec));
    return qErrorNone;
}

void QPandaSDK::setRandomSeed(int32_t seed)
{
    if (mGates) mGates->seedRandom(seed);
}

bool QPandaSDK::execute(const QRNode &n)
{
    const size_t N = (size_t)mQGatesParam->mQuantumBitNumber;
    for (size_t q : n.qubits)
        if (q >= N) return false;
    Qnum ctl(n.controls.begin(), n.controls.end());
    const int dag = n.dagger ? 1 : 0;
    HipQuantumGates &g = *mGates;
    QError r = qErrorNone;
    switch (n.kind) {
    case qsvq::HADAMARD_GATE:
        if (ctl.empty()) {
            r = g.Hadamard(n.qubits[0], 0);
        } else {
            ctl.push_back(n.qubits[0]);
            r = g.Hadamard(n.qubits[0], 0, ctl, N);
        }
        break;
    case qsvq::RESET_GATE:
        r = g.Reset(n.qubits[0]);
        break;
    case qsvq::RX_GATE:
    case qsvq::RY_GATE:
    case qsvq::RZ_GATE: {
        const size_t q = n.qubits[0];
        if (ctl.empty()) {
            r = n.kind == qsvq::RX_GATE   ? g.RX(q, n.angle, 0, dag)
                : n.kind == qsvq::RY_GATE ? g.RY(q, n.angle, 0, dag)
                                          : g.RZ(q, n.angle, 0, dag);
        } else {
            ctl.push_back(q);
            r = n.kind == qsvq::RX_GATE   ? g.RX(q, n.angle, 0, ctl, N, dag)
                : n.kind == qsvq::RY_GATE ? g.RY(q, n.angle, 0, ctl, N, dag)
                                          : g.RZ(q, n.angle, 0, ctl, N, dag);
        }
        break;
    }
    case qsvq::CNOT_GATE:
        if (ctl.empty()) {
            r = g.CNOT(n.qubits[0], n.qubits[1], 0);
        } else {
            ctl.push_back(n.qubits[0]);
            ctl.push_back(n.qubits[1]);
            r = g.CNOT(n.qubits[0], n.qubits[1], 0, ctl, (int)N);
        }
        break;
    case qsvq::NOT_GATE:
        if (ctl.empty()) {
            r = g.NOT(n.qubits[0], 0);
        } else {
            ctl.push_back(n.qubits[0]);
            r = g.NOT(n.qubits[0], 0, ctl, (int)N);
        }
        break;
    case qsvq::TOFFOLI_GATE:
        if (ctl.empty()) {
            r = g.toffoli(n.qubits[0], n.qubits[1], n.qubits[2], 0, (int)N);
        } else {                                   // Instruction.cpp:378-386: a multi-controlled NOT
            ctl.push_back(n.qubits[0]);
            ctl.push_back(n.qubits[1]);
            ctl.push_back(n.qubits[2]);
            r = g.NOT(n.qubits[2], 0, ctl, (int)N);
        }
        break;
    case qsvq::CR_GATE:                           // Instruction.cpp:517-526: controls ignored, errors swallowed
        (void)g.CR(n.qubits[0], n.qubits[1], n.angle, 0, dag);
        return true;
    case qsvq::ISWAP_GATE:                        // Instruction.cpp:558-577: the branches are swapped
        if (!ctl.empty()) r = g.iSWAP(n.qubits[0], n.qubits[1], 0, dag);
        else r = g.iSWAP(n.qubits[0], n.qubits[1], 0, ctl, (int)N, dag);
        break;
    case qsvq::SQISWAP_GATE:
        if (ctl.empty()) {
            r = g.sqiSWAP(n.qubits[0], n.qubits[1], 0, dag);
        } else {
            ctl.push_back(n.qubits[0]);
            ctl.push_back(n.qubits[1]);
            r = g.sqiSWAP(n.qubits[0], n.qubits[1], 0, ctl, (int)N, dag);
        }
        break;
    case qsvq::MEASURE_GATE: {
        const int out = g.qubitMeasure(n.qubits[0]);
        if (out == -1) return false;
        mQGatesParam->mReturnValue[n.creg] = out;
        return true;
    }
    case qsvq::PMEASURE_GATE: {
        mQGatesParam->mPMeasureSize = n.qubits.size();
        Qnum qs(n.qubits.begin(), n.qubits.end());
        r = g.pMeasure(qs, mQGatesParam->mPMeasure);
        break;
    }
    default:
        return false;                             // no instruction for this kind
    }
    return r == qErrorNone;
}

bool QPandaSDK::transmitGatesOnly(size_t count)
{
    if (!mGates || mGates->getQStateSize() == 0) return false;
    if (mGates->endGate(mQGatesParam, mGates) != qErrorNone) return false;
    for (size_t i = 0; i < count; i++)
        if (!execute(mProg.nodes[i])) return false;
    return true;
}

bool QPandaSDK::transmit() { return transmitGatesOnly(mProg.nodes.size()); }

void QPandaSDK::countState()
{
    std::stringstream ss;
    for (auto &kv : mQGatesParam->mReturnValue) ss << kv.second;
    std::string key(ss.str());
    std::reverse(key.begin(), key.end());
    auto it = mQResultMap.find(key);
    if (it != mQResultMap.end()) it->second += 1;
    else mQResultMap.insert({key, 1});
}

void QPandaSDK::countWords(const std::vector<uint64_t> &words, size_t first_measure)
{
    // registers in ascending order, each taking the bit of the last
    // measurement that writes it (as the reference's register map does)
    std::map<size_t, size_t> writer;
    for (size_t j = first_measure; j < mProg.nodes.size(); j++) writer[mProg.nodes[j].creg] = j - first_measure;
    std::vector<size_t> bitOf;
    for (auto &kv : writer) bitOf.push_back(kv.second);
    const size_t R = bitOf.size();
    std::unordered_map<uint64_t, int64_t> counts;
    for (uint64_t w : words) {
        uint64_t key = 0;
        for (size_t r = 0; r < R; r++) key |= ((w >> bitOf[r]) & 1) << r;
        counts[key] += 1;
    }
    // the string lists registers in ascending order, reversed: the highest first
    std::string s(R, '0');
    for (auto &kv : counts) {
        for (size_t p = 0; p < R; p++) s[p] = (char)('0' + ((kv.first >> (R - 1 - p)) & 1));
        mQResultMap[s] += (double)kv.second;
    }
}

void QPandaSDK::integerToBinary(size_t number, std::string &out, size_t len)
{
    out.clear();
    unsigned int index = 1u << (len - 1);
    for (size_t i = 0; i < len; i++) {
        out += (number & index) ? '1' : '0';
        index >>= 1;
    }
}

bool QPandaSDK::samplingApplies(size_t *first_measure) const
{
    const auto &ns = mProg.nodes;
    size_t fm = ns.size();
    for (size_t i = 0; i < ns.size(); i++)
        if (ns[i].kind == qsvq::MEASURE_GATE) {
            fm = i;
            break;
        }
    if (fm == ns.size() || ns.size() - fm > 64) return false;
    uint64_t seen = 0;
    for (size_t i = 0; i < ns.size(); i++) {
        const int k = ns[i].kind;
        if (k == qsvq::PMEASURE_GATE || k == qsvq::RESET_GATE) return false;
        if ((i < fm) == (k == qsvq::MEASURE_GATE)) return false;   // gates first, then measurements only
        if (k == qsvq::MEASURE_GATE) {
            const size_t q = ns[i].qubits[0];
            if (seen >> q & 1) return false;
            seen |= uint64_t(1) << q;
        }
    }
    *first_measure = fm;
    return true;
}

int QPandaSDK::runSampled(int iRepeat, size_t fm)
{
    const int32_t r0 = mGates->rngState();
    const int64_t d0 = mGates->draws();
    if (!transmitGatesOnly(fm)) {
        // as the reference's failing first run: nothing counted
        mbIsRun = false;
        miMeasureType = kPMeasureType;
        return runProgramError;
    }
    const uint64_t G = (uint64_t)(mGates->draws() - d0);
    const size_t M = mProg.nodes.size() - fm;
    std::vector<int> qubits(M);
    for (size_t j = 0; j < M; j++) qubits[j] = (int)mProg.nodes[fm + j].qubits[0];
    std::vector<uint64_t> words((size_t)iRepeat);
    int32_t rng = r0;
    int64_t ties = 0;
    if (qsv_sample_shots(mGates->handle(), qubits.data(), (int)M, (int64_t)G, &rng, iRepeat, words.data(), &ties) !=
        QSV_OK)
        return runProgramError;
    // replay the last run's measurements on the state: final state and
    // registers of the last run, and the stream position after it
    mGates->setRngState(qsv_pm_jump(r0, (uint64_t)(iRepeat - 1) * (G + M) + G));
    uint64_t last = 0;
    for (size_t j = 0; j < M; j++) {
        if (!execute(mProg.nodes[fm + j])) return runProgramError;
        last |= (uint64_t)mQGatesParam->mReturnValue[mProg.nodes[fm + j].creg] << j;
    }
    // the replayed run must be the sampler's last shot, ending where it ended
    if (mGates->rngState() != rng || last != words.back()) return runProgramError;
    countWords(words, fm);
    for (auto &kv : mQResultMap) kv.second = kv.second / iRepeat;
    miMeasureType = kMeasureType;
    mbIsRun = true;
    mSampled = true;
    return qErrorNone;
}

int QPandaSDK::run(int iRepeat)
{
    if (iRepeat <= 0) return qParameterError;
    if (!mbIsRead) return notReadFileError;
    mQGatesParam->mPMeasure.clear();
    mQGatesParam->mReturnValue.clear();
    mQResultMap.clear();
    mSampled = false;
    if (!mbIsRun) {
        if (!mGates || mQGatesParam->mQuantumBitNumber <= 0 || mGates->initState(mQGatesParam) != qErrorNone)
            return initStateError;
    }
    size_t fm = 0;
    if (!mNoSampling && samplingApplies(&fm)) return runSampled(iRepeat, fm);
    int bResult = qErrorNone;
    if (!transmit()) {
        bResult = runProgramError;
        mbIsRun = false;
    }
    if (mQGatesParam->mPMeasureSize == 0 && bResult == qErrorNone) {
        countState();
        miMeasureType = kMeasureType;
        for (int i = 1; i < iRepeat; i++) {
            if (!transmit()) {
                bResult = runProgramError;
                mbIsRun = false;
                break;
            }
            countState();
        }
        for (auto &kv : mQResultMap) kv.second = kv.second / iRepeat;
    } else {
        miMeasureType = kPMeasureType;
        std::string key;
        for (auto &p : mQGatesParam->mPMeasure) {
            integerToBinary(p.first, key, mQGatesParam->mPMeasureSize);
            mQResultMap.insert({key, p.second});
        }
    }
    if (bResult == qErrorNone) mbIsRun = true;
    return bResult;
}

std::string QPandaSDK::getResult()
{
    std::string out;
    if (!mbIsRun || mQResultMap.empty()) return out;
    for (auto &kv : mQResultMap) {
        std::stringstream ss;
        ss << kv.first << " " << kv.second << "\n";
        out += ss.str();
    }
    return out;
}

std::string QPandaSDK::getQuantumState()
{
    std::string s;
    if (!mbIsRun || !mGates) return s;
    mGates->getQState(s, mQGatesParam);
    return s;
}

QStat QPandaSDK::amplitudes() const
{
    if (!mbIsRun || !mGates) return QStat();
    return mGates->state();
}

int QPandaSDK::saveState(const std::string &path, bool binary)
{
    if (path.empty()) return qParameterError;
    if (!mbIsRun || !mGates || !mGates->handle()) return runProgramError;
    const std::string file = path + (binary ? ".fullstate.binary" : ".fullstate");
    if (qsv_save_state(mGates->handle(), file.c_str(), binary ? QSV_FILE_BINARY : QSV_FILE_TEXT, 0) != QSV_OK)
        return runProgramError;
    return qErrorNone;
}

}  // namespace QPanda

using QPanda::QPandaSDK;

namespace {
int copy_out(const std::string &s, char *buf, int *len)
{
    if (!len) return qParameterError;
    if (*len != 0 && buf) std::memcpy(buf, s.data(), std::min<size_t>((size_t)*len, s.size()));
    *len = (int)s.size();
    return qErrorNone;
}
}  // namespace

extern "C" {

int loadFileAPI(char *pcFilePath)
{
    return QPandaSDK::getIntance()->loadFile(pcFilePath ? std::string(pcFilePath) : std::string());
}
int setComputeUnitAPI(int iComputeUnit) { return QPandaSDK::getIntance()->setComputeUnit(iComputeUnit); }
int runAPI(int iRepeat) { return QPandaSDK::getIntance()->run(iRepeat); }
int getFileNameAPI(char *buf, int *pLength) { return copy_out(QPandaSDK::getIntance()->msFileName, buf, pLength); }
int getResultAPI(char *buf, int *pLength) { return copy_out(QPandaSDK::getIntance()->getResult(), buf, pLength); }
int getQuantumStateAPI(char *buf, int *pLength)
{
    return copy_out(QPandaSDK::getIntance()->getQuantumState(), buf, pLength);
}

int qrunes_parse_file(const char *path, char *buf, int *len)
{
    qsvq::QRProgram p;
    const int r = qsvq::parseQRunesFile(path ? path : "", p);
    if (len) (void)copy_out(r < 0 ? std::string() : qsvq::describe(p), buf, len);
    return r;
}

int qrunes_parse_text(const char *text, char *buf, int *len)
{
    qsvq::QRProgram p;
    const int r = qsvq::parseQRunesText(text ? text : "", p);
    if (len) (void)copy_out(r < 0 ? std::string() : qsvq::describe(p), buf, len);
    return r;
}

qpanda_sdk *qpanda_sdk_create(void) { return reinterpret_cast<qpanda_sdk *>(new QPandaSDK()); }
void qpanda_sdk_destroy(qpanda_sdk *h) { delete reinterpret_cast<QPandaSDK *>(h); }
int qpanda_sdk_load_file(qpanda_sdk *h, const char *path)
{
    return reinterpret_cast<QPandaSDK *>(h)->loadFile(path ? std::string(path) : std::string());
}
int qpanda_sdk_set_compute_unit(qpanda_sdk *h, int unit) { return reinterpret_cast<QPandaSDK *>(h)->setComputeUnit(unit); }
void qpanda_sdk_set_seed(qpanda_sdk *h, int32_t seed) { reinterpret_cast<QPandaSDK *>(h)->setRandomSeed(seed); }
void qpanda_sdk_set_resample_disabled(qpanda_sdk *h, int off)
{
    reinterpret_cast<QPandaSDK *>(h)->setResampleDisabled(off != 0);
}
int qpanda_sdk_run(qpanda_sdk *h, int repeat) { return reinterpret_cast<QPandaSDK *>(h)->run(repeat); }
int qpanda_sdk_last_run_sampled(qpanda_sdk *h) { return reinterpret_cast<QPandaSDK *>(h)->lastRunSampled(); }
int qpanda_sdk_get_result(qpanda_sdk *h, char *buf, int *len)
{
    return copy_out(reinterpret_cast<QPandaSDK *>(h)->getResult(), buf, len);
}
int qpanda_sdk_get_state(qpanda_sdk *h, char *buf, int *len)
{
    return copy_out(reinterpret_cast<QPandaSDK *>(h)->getQuantumState(), buf, len);
}
int qpanda_sdk_save_state(qpanda_sdk *h, const char *path, int binary)
{
    return reinterpret_cast<QPandaSDK *>(h)->saveState(path ? std::string(path) : std::string(), binary != 0);
}
int64_t qpanda_sdk_get_amplitudes(qpanda_sdk *h, double *out, int64_t count)
{
    const QStat st = reinterpret_cast<QPandaSDK *>(h)->amplitudes();
    const int64_t n = (int64_t)st.size();
    if (out) std::memcpy(out, st.data(), (size_t)std::min(n, std::max<int64_t>(count, 0)) * sizeof(st[0]));
    return n;
}

}  // extern "C"
