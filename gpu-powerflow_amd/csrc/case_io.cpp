// Case and matrix files (see case_io.hpp).
#include "case_io.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <tuple>

namespace cupf {
namespace {

namespace fs = std::filesystem;

bool skip_line(const std::string& line) {
    for (char ch : line) {
        if (!std::isspace(static_cast<unsigned char>(ch))) return ch == '#' || ch == '%';
    }
    return true;
}

std::ifstream open_in(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    return in;
}

std::vector<std::complex<double>> read_pairs(const std::string& path) {
    std::ifstream in = open_in(path);
    std::vector<std::complex<double>> v;
    std::string line;
    while (std::getline(in, line)) {
        if (skip_line(line)) continue;
        std::istringstream is(line);
        double re = 0.0, im = 0.0;
        if (!(is >> re >> im)) throw std::runtime_error("malformed complex entry in " + path);
        v.emplace_back(re, im);
    }
    return v;
}

std::vector<int> read_ints(const std::string& path) {
    std::ifstream in = open_in(path);
    std::vector<int> v;
    std::string line;
    while (std::getline(in, line)) {
        if (skip_line(line)) continue;
        std::istringstream is(line);
        long long x = 0;
        if (!(is >> x) || x < -(1LL << 31) || x >= (1LL << 31))
            throw std::runtime_error("malformed integer entry in " + path);
        v.push_back(int(x));
    }
    return v;
}

std::string next_payload(std::ifstream& in, const std::string& path) {
    std::string line;
    while (std::getline(in, line))
        if (!skip_line(line)) return line;
    throw std::runtime_error("unexpected end of file in " + path);
}

std::ofstream open_out(const std::string& path) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot write " + path);
    return out;
}

void write_pairs(const std::string& path, const std::vector<std::complex<double>>& v) {
    std::ofstream out = open_out(path);
    char buf[96];
    for (const auto& z : v) {
        std::snprintf(buf, sizeof(buf), "%.17e %.17e\n", z.real(), z.imag());
        out << buf;
    }
    if (!out) throw std::runtime_error("write failed: " + path);
}

void write_ints(const std::string& path, const std::vector<int>& v) {
    std::ofstream out = open_out(path);
    for (int x : v) out << x << '\n';
    if (!out) throw std::runtime_error("write failed: " + path);
}

}  // namespace

MatrixMarket read_matrix_market(const std::string& path) {
    std::ifstream in = open_in(path);
    std::string header;
    if (!std::getline(in, header)) throw std::runtime_error("missing MatrixMarket header in " + path);
    std::string h = header;
    std::transform(h.begin(), h.end(), h.begin(),
                   [](unsigned char ch) { return char(std::tolower(ch)); });
    std::istringstream hs(h);
    std::string banner, object, format, field, symmetry;
    hs >> banner >> object >> format >> field >> symmetry;
    if (banner != "%%matrixmarket" || object != "matrix")
        throw std::runtime_error("not a MatrixMarket matrix: " + path);
    MatrixMarket m;
    if (field == "complex") m.is_complex = true;
    else if (field != "real" && field != "integer")
        throw std::runtime_error("unsupported MatrixMarket field '" + field + "' in " + path);
    if (symmetry != "general" && symmetry != "symmetric")
        throw std::runtime_error("unsupported MatrixMarket symmetry '" + symmetry + "' in " + path);
    const bool sym = symmetry == "symmetric";
    const int w = m.is_complex ? 2 : 1;
    std::istringstream ds(next_payload(in, path));
    if (format == "array") {
        if (sym) throw std::runtime_error("symmetric array files are not supported: " + path);
        if (!(ds >> m.rows >> m.cols) || m.rows < 0 || m.cols < 0)
            throw std::runtime_error("malformed MatrixMarket size line in " + path);
        m.is_dense = true;
        m.values.resize(size_t(m.rows) * m.cols * w);
        for (size_t k = 0; k < size_t(m.rows) * m.cols; ++k) {
            std::istringstream es(next_payload(in, path));
            for (int c = 0; c < w; ++c)
                if (!(es >> m.values[k * w + c]))
                    throw std::runtime_error("malformed MatrixMarket array entry in " + path);
        }
        return m;
    }
    if (format != "coordinate")
        throw std::runtime_error("unsupported MatrixMarket format '" + format + "' in " + path);
    long long nnz = 0;
    if (!(ds >> m.rows >> m.cols >> nnz) || m.rows < 0 || m.cols < 0 || nnz < 0)
        throw std::runtime_error("malformed MatrixMarket size line in " + path);
    struct E { int r, c; double re, im; };
    std::vector<E> ent;
    ent.reserve(size_t(sym ? 2 * nnz : nnz));
    for (long long k = 0; k < nnz; ++k) {
        std::istringstream es(next_payload(in, path));
        long long r = 0, c = 0;
        double re = 0.0, im = 0.0;
        if (!(es >> r >> c >> re) || (m.is_complex && !(es >> im)))
            throw std::runtime_error("malformed MatrixMarket entry in " + path);
        if (r < 1 || r > m.rows || c < 1 || c > m.cols)
            throw std::runtime_error("MatrixMarket entry out of range in " + path);
        ent.push_back({int(r - 1), int(c - 1), re, im});
        if (sym && r != c) ent.push_back({int(c - 1), int(r - 1), re, im});
    }
    std::stable_sort(ent.begin(), ent.end(),
                     [](const E& a, const E& b) { return std::tie(a.r, a.c) < std::tie(b.r, b.c); });
    m.indptr.assign(size_t(m.rows) + 1, 0);
    for (size_t k = 0; k < ent.size(); ++k) {
        if (k > 0 && ent[k].r == ent[k - 1].r && ent[k].c == ent[k - 1].c) {
            m.values[m.values.size() - w] += ent[k].re;          // duplicates are summed
            if (m.is_complex) m.values.back() += ent[k].im;
            continue;
        }
        ++m.indptr[size_t(ent[k].r) + 1];
        m.indices.push_back(ent[k].c);
        m.values.push_back(ent[k].re);
        if (m.is_complex) m.values.push_back(ent[k].im);
    }
    for (int r = 0; r < m.rows; ++r) m.indptr[size_t(r) + 1] += m.indptr[size_t(r)];
    return m;
}

void write_matrix_market_csr(const std::string& path, int rows, int cols, const int* indptr,
                             const int* indices, const double* values, bool is_complex) {
    if (rows < 0 || cols < 0 || !indptr || (indptr[rows] > 0 && (!indices || !values)))
        throw std::invalid_argument("write_matrix_market_csr: bad matrix");
    std::ofstream out = open_out(path);
    out << "%%MatrixMarket matrix coordinate " << (is_complex ? "complex" : "real") << " general\n";
    out << rows << ' ' << cols << ' ' << indptr[rows] << '\n';
    char buf[128];
    for (int r = 0; r < rows; ++r)
        for (int k = indptr[r]; k < indptr[r + 1]; ++k) {
            if (is_complex)
                std::snprintf(buf, sizeof(buf), "%d %d %.17e %.17e\n", r + 1, indices[k] + 1,
                              values[2 * k], values[2 * k + 1]);
            else
                std::snprintf(buf, sizeof(buf), "%d %d %.17e\n", r + 1, indices[k] + 1, values[k]);
            out << buf;
        }
    if (!out) throw std::runtime_error("write failed: " + path);
}

void write_matrix_market_dense(const std::string& path, int rows, int cols, const double* values) {
    if (rows < 0 || cols < 0 || (size_t(rows) * cols > 0 && !values))
        throw std::invalid_argument("write_matrix_market_dense: bad array");
    std::ofstream out = open_out(path);
    out << "%%MatrixMarket matrix array real general\n" << rows << ' ' << cols << '\n';
    char buf[64];
    for (size_t k = 0; k < size_t(rows) * cols; ++k) {
        std::snprintf(buf, sizeof(buf), "%.17e\n", values[k]);
        out << buf;
    }
    if (!out) throw std::runtime_error("write failed: " + path);
}

DumpCase load_dump_case(const std::string& dir) {
    const fs::path d(dir);
    if (!fs::is_directory(d)) throw std::runtime_error("case directory does not exist: " + dir);
    DumpCase c;
    c.name = fs::absolute(d).lexically_normal().filename().string();
    if (c.name.empty()) c.name = fs::absolute(d).lexically_normal().parent_path().filename().string();
    const MatrixMarket y = read_matrix_market((d / "dump_Ybus.mtx").string());
    if (y.is_dense || !y.is_complex)
        throw std::runtime_error("dump_Ybus.mtx must be a complex coordinate matrix");
    c.rows = y.rows;
    c.cols = y.cols;
    c.indptr = y.indptr;
    c.indices = y.indices;
    c.ybus.resize(y.indices.size());
    for (size_t k = 0; k < c.ybus.size(); ++k) c.ybus[k] = {y.values[2 * k], y.values[2 * k + 1]};
    c.sbus = read_pairs((d / "dump_Sbus.txt").string());
    c.v0 = read_pairs((d / "dump_V.txt").string());
    c.pv = read_ints((d / "dump_pv.txt").string());
    c.pq = read_ints((d / "dump_pq.txt").string());
    if (fs::exists(d / "dump_Vref.txt")) c.vref = read_pairs((d / "dump_Vref.txt").string());
    if (c.rows <= 0 || c.rows != c.cols) throw std::runtime_error("Ybus must be square and non-empty");
    if (int(c.sbus.size()) != c.rows) throw std::runtime_error("Sbus size does not match Ybus rows");
    if (int(c.v0.size()) != c.rows) throw std::runtime_error("V0 size does not match Ybus rows");
    if (!c.vref.empty() && int(c.vref.size()) != c.rows)
        throw std::runtime_error("Vref size does not match Ybus rows");
    for (const auto* v : {&c.pv, &c.pq})
        for (int b : *v)
            if (b < 0 || b >= c.rows) throw std::runtime_error("bus index out of range in pv/pq");
    return c;
}

void write_dump_case(const std::string& dir, const DumpCase& c) {
    if (c.rows <= 0 || c.rows != c.cols || int(c.indptr.size()) != c.rows + 1 ||
        c.indices.size() != c.ybus.size() || int(c.sbus.size()) != c.rows ||
        int(c.v0.size()) != c.rows || (!c.vref.empty() && int(c.vref.size()) != c.rows))
        throw std::invalid_argument("write_dump_case: inconsistent case");
    const fs::path d(dir);
    fs::create_directories(d);
    std::vector<double> y(2 * c.ybus.size());
    for (size_t k = 0; k < c.ybus.size(); ++k) {
        y[2 * k] = c.ybus[k].real();
        y[2 * k + 1] = c.ybus[k].imag();
    }
    write_matrix_market_csr((d / "dump_Ybus.mtx").string(), c.rows, c.cols, c.indptr.data(),
                            c.indices.data(), y.data(), true);
    write_pairs((d / "dump_Sbus.txt").string(), c.sbus);
    write_pairs((d / "dump_V.txt").string(), c.v0);
    write_ints((d / "dump_pv.txt").string(), c.pv);
    write_ints((d / "dump_pq.txt").string(), c.pq);
    if (!c.vref.empty()) write_pairs((d / "dump_Vref.txt").string(), c.vref);
}

}  // namespace cupf
