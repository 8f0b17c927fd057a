// Case and matrix files: the reference's dump-case directories and its
// MatrixMarket Jacobian / right-hand-side files.
//
// Dump case (read by cuPF/tests/dump_case_loader.cpp:92-231, written by
// benchmark/common/matpower_data.py:350-358): a directory with
//   dump_Ybus.mtx  MatrixMarket coordinate complex (general or symmetric,
//                  1-based, duplicates summed), the bus admittance matrix;
//   dump_Sbus.txt  one "re im" pair per line, the net injections;
//   dump_V.txt     one "re im" pair per line, the initial voltages;
//   dump_pv.txt    one 0-based bus index per line (PV buses);
//   dump_pq.txt    one 0-based bus index per line (PQ buses);
//   dump_Vref.txt  optional, "re im" per line, a reference solution.
// Blank lines and lines starting with '#' or '%' are skipped.
//
// Linear-system datasets (custom_linear_solver/tests/datasets/README.md:1-20):
// J.mtx "coordinate real general" and F.mtx "array real general" (n x 1).
#pragma once

#include <complex>
#include <string>
#include <vector>

namespace cupf {

struct DumpCase {
    std::string name;
    int rows = 0, cols = 0;
    std::vector<int> indptr, indices;                 // Ybus CSR, columns sorted per row
    std::vector<std::complex<double>> ybus, sbus, v0, vref;   // vref empty when absent
    std::vector<int> pv, pq;
};

// Loads and checks a dump-case directory (square Ybus, Sbus and V0 of its
// size, bus indices in range); throws std::runtime_error with the file name
// on any malformed or missing input.
DumpCase load_dump_case(const std::string& dir);
// Writes the directory (created if missing) in the same format; the Ybus as
// "coordinate complex general", 17 significant digits.
void write_dump_case(const std::string& dir, const DumpCase& c);

// A MatrixMarket file: a coordinate matrix as CSR (columns sorted per row,
// duplicates summed, the mirror of every off-diagonal entry of a symmetric
// file added), or a dense array (column-major).
struct MatrixMarket {
    int rows = 0, cols = 0;
    bool is_complex = false, is_dense = false;
    std::vector<int> indptr, indices;
    std::vector<double> values;                       // real, or interleaved (re, im)
};

MatrixMarket read_matrix_market(const std::string& path);
void write_matrix_market_csr(const std::string& path, int rows, int cols, const int* indptr,
                             const int* indices, const double* values, bool is_complex);
void write_matrix_market_dense(const std::string& path, int rows, int cols, const double* values);

}  // namespace cupf
