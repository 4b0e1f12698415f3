// QuantumGates.cpp -- base-class members of the gate-API interface.
#include "QuantumGates.h"

#include <ctime>

QuantumGates::QuantumGates() {}

QuantumGates::~QuantumGates() {}

// Same generator, seed expression and process-wide state as
// QPandaSDK/QuantumInstructionHandle/QuantumGates.cpp:46-85: the 16807
// multiplier with Schrage's decomposition, seeded on the first call from the
// local time.
double QuantumGates::randGenerator()
{
    const int ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
    static int state = [] {
        time_t raw;
        time(&raw);
        const struct tm *t = localtime(&raw);
        return t->tm_year + 70 * (t->tm_mon + 1 + 12 * (t->tm_mday + 31 * (t->tm_hour + 23 * (t->tm_min + 59 * t->tm_sec))));
    }();
    int next = ia * (state % iq) - ir * (state / iq);
    if (next < 0) next += im;
    state = next;
    return (double)next / im;
}
