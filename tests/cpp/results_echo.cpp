// results_echo.cpp -- writes a qsl::hipbench::Table read from stdin, so a
// test can hold the C++ writer against a file the reference's own
// Results::writeToFile wrote (tests/golden/np_phase.csv).  No device work.
//
// stdin: <nmeta>\n <meta lines>\n <headings line>\n <nint> <nrows>\n
//        then nrows rows of values separated by spaces.
#include <qsl/benchmark/hip_results.hpp>

#include <iostream>
#include <sstream>
#include <string>
#include <vector>

int main()
{
    std::string line;
    std::getline(std::cin, line);
    const int nmeta = std::stoi(line);
    std::vector<std::string> meta(nmeta);
    for (auto &m : meta) std::getline(std::cin, m);
    std::getline(std::cin, line);
    std::istringstream hs(line);
    std::vector<std::string> head;
    for (std::string h; hs >> h;) head.push_back(h);
    std::size_t nint = 0, nrows = 0;
    std::cin >> nint >> nrows;
    qsl::hipbench::Table<unsigned, double> t(head);
    for (const auto &m : meta) t.addMeta(m);
    for (std::size_t r = 0; r < nrows; r++) {
        std::vector<unsigned> ints(nint);
        std::vector<double> fps(head.size() - nint);
        for (auto &v : ints) std::cin >> v;
        for (auto &v : fps) std::cin >> v;
        t.addRow(ints, fps);
    }
    t.write(std::cout);
    return 0;
}
