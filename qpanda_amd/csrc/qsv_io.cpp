// qsv_io.cpp -- state I/O: whole-state host copies and state files, streamed
// through two pinned staging buffers so that the copy of one chunk overlaps
// the conversion / file access of the previous one and no host copy of the
// whole state is made for complex64 states.
//
// Files (QCommand.cpp:402-487 writes both from the state text):
//   binary  2^n records {f64 re, f64 im} in index order (.fullstate.binary);
//   text    "state[i].real = X state[i].imag = Y\n" per amplitude, the
//           getQState text (CPUQuantumGates.cpp:64-82; .fullstate).
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "qsv_internal.h"

namespace qsv {
namespace {

constexpr uint64_t kChunkAmps = uint64_t(1) << 22;   // 4 Mi amplitudes: 64 MiB of f64 records

struct Stage {
    hipStream_t stream = nullptr;
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<double> conv;      // f64 records of a complex64 chunk

    ~Stage()
    {
        // copies in flight read or write the pinned buffers: finish them first
        if (stream) (void)hipStreamSynchronize(stream);
        for (int i = 0; i < 2; i++) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (pin[i]) (void)hipHostFree(pin[i]);
        }
    }
};

int make_stage(qsv_state *s, Stage &st)
{
    const uint64_t chunk = std::min<uint64_t>(kChunkAmps, s->dim);
    const size_t bytes = (size_t)chunk * (s->prec == QSV_C128 ? 16 : 8);
    st.stream = s->stream;
    for (int i = 0; i < 2; i++) {
        QSV_HIP(hipHostMalloc(&st.pin[i], bytes, hipHostMallocDefault));
        QSV_HIP(hipEventCreateWithFlags(&st.ev[i], hipEventDisableTiming));
    }
    if (s->prec == QSV_C64) st.conv.resize((size_t)chunk * 2);
    return QSV_OK;
}

// Streams the state out as f64 records: sink(first, count, records) for each
// chunk in index order.
int stream_out(qsv_state *s, const std::function<int(uint64_t, uint64_t, const double *)> &sink)
{
    Stage st;
    int rc = make_stage(s, st);
    if (rc) return rc;
    const size_t esz = s->prec == QSV_C128 ? 16 : 8;
    const uint64_t nchunks = (s->dim + kChunkAmps - 1) / kChunkAmps;
    auto enqueue = [&](uint64_t c) -> int {
        const uint64_t first = c * kChunkAmps, cnt = std::min(kChunkAmps, s->dim - first);
        QSV_HIP(hipMemcpyAsync(st.pin[c & 1], (const char *)s->amps + first * esz, cnt * esz,
                               hipMemcpyDeviceToHost, s->stream));
        QSV_HIP(hipEventRecord(st.ev[c & 1], s->stream));
        return QSV_OK;
    };
    if ((rc = enqueue(0))) return rc;
    for (uint64_t c = 0; c < nchunks; c++) {
        if (c + 1 < nchunks && (rc = enqueue(c + 1))) return rc;
        QSV_HIP(hipEventSynchronize(st.ev[c & 1]));
        const uint64_t first = c * kChunkAmps, cnt = std::min(kChunkAmps, s->dim - first);
        const double *rec = (const double *)st.pin[c & 1];
        if (s->prec == QSV_C64) {
            const float *f = (const float *)st.pin[c & 1];
            for (uint64_t i = 0; i < 2 * cnt; i++) st.conv[i] = f[i];
            rec = st.conv.data();
        }
        if ((rc = sink(first, cnt, rec))) return rc;
    }
    return QSV_OK;
}

// Streams f64 records into the state: source(first, count, records) fills
// each chunk in index order.
int stream_in(qsv_state *s, const std::function<int(uint64_t, uint64_t, double *)> &source)
{
    Stage st;
    int rc = make_stage(s, st);
    if (rc) return rc;
    const size_t esz = s->prec == QSV_C128 ? 16 : 8;
    const uint64_t nchunks = (s->dim + kChunkAmps - 1) / kChunkAmps;
    for (uint64_t c = 0; c < nchunks; c++) {
        const int slot = (int)(c & 1);
        if (c >= 2) QSV_HIP(hipEventSynchronize(st.ev[slot]));     // the slot's last copy is done
        const uint64_t first = c * kChunkAmps, cnt = std::min(kChunkAmps, s->dim - first);
        double *rec = s->prec == QSV_C128 ? (double *)st.pin[slot] : st.conv.data();
        if ((rc = source(first, cnt, rec))) return rc;
        if (s->prec == QSV_C64) {
            float *f = (float *)st.pin[slot];
            for (uint64_t i = 0; i < 2 * cnt; i++) f[i] = (float)rec[i];
        }
        QSV_HIP(hipMemcpyAsync((char *)s->amps + first * esz, st.pin[slot], cnt * esz, hipMemcpyHostToDevice,
                               s->stream));
        QSV_HIP(hipEventRecord(st.ev[slot], s->stream));
    }
    QSV_HIP(hipStreamSynchronize(s->stream));
    return QSV_OK;
}

}  // namespace

// qubitMeasure's P(0) in the CPU simulator's own order
// (CPUQuantumGates.cpp:1074-1083): every index with bit q = 0 in ascending
// order, p += |a| |a| with |a| = hypot(re, im) from the C library (what
// std::abs of a std::complex<double> computes there), each operation rounded
// separately (no contraction into an FMA).  The state streams through the
// pinned chunks of stream_out, the copy of the next chunk overlapping the sum
// of this one.  Used only inside the measurement tie band.
int serial_prob0_host(qsv_state *s, int q, double *p0)
{
    double p = 0;
    const uint64_t bit = uint64_t(1) << q;
    const int rc = stream_out(s, [&](uint64_t first, uint64_t cnt, const double *rec) {
#pragma clang fp contract(off)
        for (uint64_t k = 0; k < cnt; k++) {
            if ((first + k) & bit) continue;
            const double h = std::hypot(rec[2 * k], rec[2 * k + 1]);
            p = p + h * h;
        }
        return QSV_OK;
    });
    if (rc) return rc;
    *p0 = p;
    return QSV_OK;
}

namespace {

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

// One getQState line per amplitude.  digits 0: std::ostream's default (6
// significant digits, %g); otherwise %.<digits>g.
size_t format_line(char *out, size_t cap, uint64_t i, double re, double im, int digits)
{
    const int d = digits > 0 ? digits : 6;
    const int len = std::snprintf(out, cap, "state[%" PRIu64 "].real = %.*g state[%" PRIu64 "].imag = %.*g\n", i, d,
                                  re, i, d, im);
    return len > 0 ? (size_t)len : 0;
}

// Parses "state[<i>].real = <x> state[<i>].imag = <y>" at p; returns the end or nullptr.
const char *parse_line(const char *p, uint64_t want, double *re, double *im)
{
    for (int part = 0; part < 2; part++) {
        while (*p == ' ' || *p == '\t') p++;
        if (std::strncmp(p, "state[", 6) != 0) return nullptr;
        char *end = nullptr;
        errno = 0;
        const unsigned long long idx = std::strtoull(p + 6, &end, 10);
        if (errno || end == p + 6 || idx != want) return nullptr;
        p = end;
        const char *field = part == 0 ? "].real = " : "].imag = ";
        if (std::strncmp(p, field, 9) != 0) return nullptr;
        p += 9;
        const double v = std::strtod(p, &end);
        if (end == p) return nullptr;
        (part == 0 ? *re : *im) = v;
        p = end;
    }
    while (*p == ' ' || *p == '\t' || *p == '\r') p++;
    if (*p != '\n' && *p != '\0') return nullptr;
    return *p == '\n' ? p + 1 : p;
}

}  // namespace
}  // namespace qsv

using namespace qsv;

#define CHECK_STATE(s) QSV_CHECK_STATE(s)

int qsv_set_state(qsv_state *s, const double *amps)
{
    CHECK_STATE(s);
    if (!amps) return fail(QSV_ERR_ARG, "null amps");
    if (s->prec == QSV_C128) {
        QSV_HIP(hipMemcpyAsync(s->amps, amps, s->dim * 16, hipMemcpyHostToDevice, s->stream));
        QSV_HIP(hipStreamSynchronize(s->stream));
        return QSV_OK;
    }
    return stream_in(s, [&](uint64_t first, uint64_t cnt, double *rec) {
        std::memcpy(rec, amps + 2 * first, cnt * 16);
        return QSV_OK;
    });
}

int qsv_get_state(qsv_state *s, double *amps)
{
    CHECK_STATE(s);
    if (!amps) return fail(QSV_ERR_ARG, "null amps");
    if (s->prec == QSV_C128) {
        QSV_HIP(hipMemcpyAsync(amps, s->amps, s->dim * 16, hipMemcpyDeviceToHost, s->stream));
        QSV_HIP(hipStreamSynchronize(s->stream));
        return QSV_OK;
    }
    return stream_out(s, [&](uint64_t first, uint64_t cnt, const double *rec) {
        std::memcpy(amps + 2 * first, rec, cnt * 16);
        return QSV_OK;
    });
}

int qsv_save_state(qsv_state *s, const char *path, int format, int digits)
{
    CHECK_STATE(s);
    if (!path || !*path) return fail(QSV_ERR_ARG, "empty path");
    if (format != QSV_FILE_BINARY && format != QSV_FILE_TEXT) return fail(QSV_ERR_ARG, "unknown file format");
    if (digits < 0 || digits > 17) return fail(QSV_ERR_ARG, "digits must be 0..17");
    File f;
    f.f = std::fopen(path, format == QSV_FILE_BINARY ? "wb" : "w");
    if (!f.f) return fail(QSV_ERR_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
    std::vector<char> text;
    int rc = stream_out(s, [&](uint64_t first, uint64_t cnt, const double *rec) {
        if (format == QSV_FILE_BINARY) {
            if (std::fwrite(rec, 16, cnt, f.f) != cnt) return fail(QSV_ERR_ARG, "write failed");
            return QSV_OK;
        }
        text.resize((size_t)cnt * 96);
        size_t len = 0;
        for (uint64_t i = 0; i < cnt; i++)
            len += format_line(text.data() + len, text.size() - len, first + i, rec[2 * i], rec[2 * i + 1], digits);
        if (std::fwrite(text.data(), 1, len, f.f) != len) return fail(QSV_ERR_ARG, "write failed");
        return QSV_OK;
    });
    if (rc) return rc;
    if (std::fflush(f.f) != 0) return fail(QSV_ERR_ARG, "write failed");
    return QSV_OK;
}

int qsv_load_state(qsv_state *s, const char *path, int format)
{
    CHECK_STATE(s);
    if (!path || !*path) return fail(QSV_ERR_ARG, "empty path");
    if (format != QSV_FILE_BINARY && format != QSV_FILE_TEXT) return fail(QSV_ERR_ARG, "unknown file format");
    File f;
    f.f = std::fopen(path, format == QSV_FILE_BINARY ? "rb" : "r");
    if (!f.f) return fail(QSV_ERR_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
    if (format == QSV_FILE_BINARY) {
        if (std::fseek(f.f, 0, SEEK_END) != 0) return fail(QSV_ERR_ARG, "cannot size the file");
        const long long size = std::ftell(f.f);
        if (size < 0 || (unsigned long long)size != s->dim * 16)
            return fail(QSV_ERR_ARG, "binary state file holds " + std::to_string(size) + " bytes, the state needs " +
                                         std::to_string(s->dim * 16));
        std::rewind(f.f);
        return stream_in(s, [&](uint64_t, uint64_t cnt, double *rec) {
            if (std::fread(rec, 16, cnt, f.f) != cnt) return fail(QSV_ERR_ARG, "short read");
            return QSV_OK;
        });
    }
    std::vector<char> line(4096);
    int rc = stream_in(s, [&](uint64_t first, uint64_t cnt, double *rec) {
        for (uint64_t i = 0; i < cnt; i++) {
            if (!std::fgets(line.data(), (int)line.size(), f.f))
                return fail(QSV_ERR_ARG, "state text ends at amplitude " + std::to_string(first + i));
            if (!parse_line(line.data(), first + i, &rec[2 * i], &rec[2 * i + 1]))
                return fail(QSV_ERR_ARG, "bad state text at amplitude " + std::to_string(first + i));
        }
        return QSV_OK;
    });
    if (rc) return rc;
    // nothing but blank lines may follow
    while (std::fgets(line.data(), (int)line.size(), f.f))
        for (const char *p = line.data(); *p; p++)
            if (*p != ' ' && *p != '\n' && *p != '\r' && *p != '\t') return fail(QSV_ERR_ARG, "state text is longer than the state");
    return QSV_OK;
}
