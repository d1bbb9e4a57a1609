// qsl/benchmark/hip_results.hpp -- a table of timing results in qsl's
// on-disk results format, so the MI355X backend's sweeps sit beside the
// reference's own result files (results/np_phase.csv) and are read by the
// same scripts.
//
// The format is the one Results<T,R>::writeToFile writes
// (include/qsl/benchmark/results.tpp:316-355):
//   "# <meta line>"              one per metadata line
//   ""                           one empty line
//   "<head0> <head1> ..."        column headings, space separated
//   "<int cols> <fp cols>"       one row per line, integer columns first,
//                                values in the default std::ostream format
// addRow refuses rows whose column counts disagree with the headings (first
// row) or with the first row, with std::logic_error, as Results::addRow does
// (results.tpp:65-98).
//
// The class is qsl::hipbench::Table, not qsl::Results, so that it can be
// used next to the reference's headers in one program.
#ifndef QSL_BENCHMARK_HIP_RESULTS_HPP
#define QSL_BENCHMARK_HIP_RESULTS_HPP

#include <cmath>
#include <concepts>
#include <fstream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace qsl::hipbench {

template <std::integral T, std::floating_point R>
class Table {
public:
    explicit Table(std::vector<std::string> headings) : head_(std::move(headings)) {}

    void addMeta(const std::string &line) { meta_.push_back(line); }

    void addRow(const std::vector<T> &ints, const std::vector<R> &fps)
    {
        if (ints_.empty()) {
            if (ints.size() + fps.size() != head_.size())
                throw std::logic_error("Total number of columns does not match number of "
                                       "column headings in first call to addRow");
        } else if (ints.size() != ints_.front().size()) {
            throw std::logic_error("Incorrect number of integer columns in call to addRow");
        } else if (fps.size() != fps_.front().size()) {
            throw std::logic_error("Incorrect number of floating point columns in call to addRow");
        }
        ints_.push_back(ints);
        fps_.push_back(fps);
    }

    const std::vector<std::string> &headings() const { return head_; }
    const std::vector<std::string> &meta() const { return meta_; }
    std::size_t rows() const { return ints_.size(); }

    // column `c` (integer columns first) as doubles
    std::vector<double> column(std::size_t c) const
    {
        if (c >= head_.size()) throw std::out_of_range("column index out of range");
        std::vector<double> v;
        for (std::size_t r = 0; r < ints_.size(); r++) {
            const std::size_t ni = ints_[r].size();
            v.push_back(c < ni ? (double)ints_[r][c] : (double)fps_[r][c - ni]);
        }
        return v;
    }

    double mean(std::size_t c) const
    {
        const auto v = column(c);
        double s = 0;
        for (double x : v) s += x;
        return s / (double)v.size();
    }

    void write(std::ostream &os) const
    {
        for (const auto &m : meta_) os << "# " << m << '\n';
        os << '\n';
        for (std::size_t i = 0; i < head_.size(); i++)
            os << head_[i] << (i + 1 < head_.size() ? " " : "\n");
        for (std::size_t r = 0; r < ints_.size(); r++) {
            // integer columns, then floating-point columns; the last value
            // ends the line
            const std::size_t ni = ints_[r].size(), nf = fps_[r].size();
            for (std::size_t i = 0; i < ni; i++) os << ints_[r][i] << (i + 1 < ni + nf ? " " : "\n");
            for (std::size_t i = 0; i < nf; i++) os << fps_[r][i] << (i + 1 < nf ? " " : "\n");
        }
    }

    void writeToFile(const std::string &filename) const
    {
        std::ofstream f(filename);
        if (!f) throw std::runtime_error("cannot open " + filename);
        write(f);
    }

private:
    std::vector<std::string> head_, meta_;
    std::vector<std::vector<T>> ints_;
    std::vector<std::vector<R>> fps_;
};

}  // namespace qsl::hipbench

#endif
