// QRunes.h -- QRunes program front end for the MI355X engine.
//
// Reads a QRunes text program into a flat instruction list with the
// reference front end's semantics (QPandaSDK/QRunesParser: the lexer's token
// rules, lex.yy.cpp:914-1079; the grammar and its actions, y.tab.cpp:1290-1650;
// the DAGGER / CONTROL list controller, LinkedListController.cpp:51-339):
//   * keywords QINIT CREG H HADAMARD NOT RESET RX RY RZ ISWAP CR CPHASE SQISWAP
//     CNOT TOFFOLI MEASURE PMEASURE CONTROL ENDCONTROL DAGGER ENDDAGGER, case
//     sensitive, longest match; "pi" in any letter case; integers and d.d
//     decimals; '%' starts a line comment; any other character is skipped;
//   * program = QINIT n, CREG m, one or more statements, optionally one final
//     PMEASURE list; ISWAP and CR / CPHASE take two qubits and a quoted angle;
//   * angles: + - * / ^ and parentheses, '^' left-associative and tighter than
//     '*' '/', unary +/- tightest; "a - b" evaluates to a + b (the grammar's own
//     action); a quotient or power that is not a normal double is an error;
//   * qubit indices must be < n, classical registers < m;
//   * DAGGER blocks reverse their statements and mark them daggered, nested
//     blocks by the controller's parity rule; CONTROL q adds q to the control
//     list of every following gate until ENDCONTROL q; MEASURE / PMEASURE
//     inside DAGGER or CONTROL are dropped, as are the statements of a DAGGER
//     left open; an ENDDAGGER / ENDCONTROL without its opener and a repeated
//     CONTROL are ignored.
#ifndef QSV_QRUNES_H
#define QSV_QRUNES_H

#include <cstddef>
#include <string>
#include <vector>

namespace qsvq {

// Gate specifiers, numbered as the reference's enum (QParamStruct.h:10-22).
enum GateKind {
    HADAMARD_GATE = 1,
    RX_GATE,
    RY_GATE,
    RZ_GATE,
    NOT_GATE,
    TOFFOLI_GATE,
    RESET_GATE,
    CR_GATE,
    ISWAP_GATE,
    SQISWAP_GATE,
    CNOT_GATE,
    MEASURE_GATE,
    PMEASURE_GATE
};

struct QRNode {
    int kind = 0;                      // GateKind
    std::vector<size_t> qubits;        // operands in program order
    double angle = 0;                  // RX / RY / RZ / CR / ISWAP
    bool dagger = false;
    std::vector<size_t> controls;      // CONTROL list in effect (gates only)
    size_t creg = 0;                   // MEASURE
};

struct QRProgram {
    int qubits = 0;                    // QINIT
    int cregs = 0;                     // CREG
    std::vector<QRNode> nodes;
};

// Parse program text.  Returns the QINIT qubit count, or -1 on any syntax or
// semantic error (the reference's QRunesParser returns -1 in both cases).
int parseQRunesText(const std::string &text, QRProgram &prog, std::string *error = nullptr);

// Parse a program file; -1 also when the file cannot be read
// (QRunesParserAPI.cpp:28-44).
int parseQRunesFile(const std::string &path, QRProgram &prog, std::string *error = nullptr);

// One line per node: "<kind> q0 q1 ... | angle=<%.17g> dagger=<0|1> ctrl=a,b creg=c"
std::string describe(const QRProgram &prog);

}  // namespace qsvq

#endif
