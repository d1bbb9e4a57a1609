// sweep_hip.cpp -- the gate-timing sweep (qsl/benchmark/hip_sweep.hpp) of
// Qubits<Hip, double> and Qubits<Hip, float>, written in qsl's results
// format.  Needs an MI355X.
//
//   sweep_hip <gate> <out file> [len_factor] [nmin] [nmax]
#include <qsl/qubits.hpp>
#include <qsl/benchmark/hip_sweep.hpp>

#include <cstdlib>
#include <cstring>
#include <iostream>

int main(int argc, char **argv)
{
    using namespace qsl::hipbench;
    if (argc < 3) {
        std::cerr << "usage: sweep_hip <gate> <out file> [len_factor] [nmin] [nmax]\n";
        return 2;
    }
    const Gate all[] = {Gate::pauliX, Gate::hadamard, Gate::rotateX, Gate::rotateY,
                        Gate::rotateZ, Gate::phase, Gate::controlNot, Gate::controlPhase};
    Gate g = Gate::rotateX;
    bool found = false;
    for (Gate x : all)
        if (std::strcmp(argv[1], gateName(x)) == 0) {
            g = x;
            found = true;
        }
    if (!found) {
        std::cerr << "unknown gate " << argv[1] << "\n";
        return 2;
    }
    const unsigned lf = argc > 3 ? (unsigned)std::atoi(argv[3]) : 20;
    const unsigned nmin = argc > 4 ? (unsigned)std::atoi(argv[4]) : 2;
    const unsigned nmax = argc > 5 ? (unsigned)std::atoi(argv[5]) : 19;
    nqubitSweep<qsl::Qubits<qsl::Type::Hip, double>, qsl::Qubits<qsl::Type::Hip, float>>(
        g, argv[2], lf, nmin, nmax);
    return 0;
}
