// Host matrix arrays and the CSV format of the reference (src/matrix.hpp:93-188,410-646):
// first line "# rows,cols[,count]", then rows*count lines of comma-separated values, the
// matrices stacked back to back.  Boost-free.
#pragma once

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace xcli {

template <typename T>
struct matrix_array {
    uint32_t rows = 0, cols = 0, count = 0;
    std::vector<T> data;

    matrix_array() = default;
    matrix_array(uint32_t r, uint32_t c, uint32_t n) : rows(r), cols(c), count(n), data((size_t)r * c * n) {}
    size_t area() const { return (size_t)rows * cols; }
    size_t size() const { return data.size(); }
    T* matrix(uint32_t i) { return data.data() + area() * i; }
    const T* matrix(uint32_t i) const { return data.data() + area() * i; }
};

// read_header, src/matrix.hpp:94-127: "# rows,cols[,count]".
inline void read_header(std::istream& in, uint32_t& rows, uint32_t& cols, uint32_t& count) {
    std::string line;
    if (!std::getline(in, line)) throw std::runtime_error("Could not read file header");
    if (line.compare(0, 2, "# ") != 0) throw std::runtime_error("Invalid file header format");
    std::vector<std::string> tok;
    std::stringstream ss(line.substr(2));
    std::string t;
    while (std::getline(ss, t, ',')) tok.push_back(t);
    if (tok.empty() || tok[0].find_first_not_of(" \t\r") == std::string::npos)
        throw std::runtime_error("Missing matrix header");
    if (tok.size() < 2) throw std::runtime_error("Missing number of columns");
    if (tok.size() > 3) throw std::runtime_error("Invalid matrix header format, too many tokens");
    auto num = [](const std::string& s) -> uint32_t {
        size_t pos = 0;
        long v = std::stol(s, &pos);
        if (v < 0) throw std::runtime_error("Negative matrix dimension in header");
        return (uint32_t)v;
    };
    rows = num(tok[0]);
    cols = num(tok[1]);
    count = tok.size() == 3 ? num(tok[2]) : 1;
}

template <typename T>
T parse_value(const std::string& s) {
    size_t pos = 0;
    double v = std::stod(s, &pos);
    return (T)v;
}

// read_data, src/matrix.hpp:129-168: one matrix of rows lines; a short or long line throws.
template <typename T>
void read_data(std::istream& in, uint32_t rows, uint32_t cols, T* out) {
    std::string line;
    for (uint32_t y = 0; y < rows; ++y) {
        if (!std::getline(in, line))
            throw std::runtime_error("Failed to read line " + std::to_string(y) + " from csv file");
        if (!line.empty() && line.back() == '\r') line.pop_back();
        uint32_t x = 0;
        size_t start = 0;
        while (start <= line.size()) {
            size_t end = line.find(',', start);
            if (end == std::string::npos) end = line.size();
            if (x >= cols)
                throw std::runtime_error("Line " + std::to_string(y) + " is too long, expected " +
                                         std::to_string(cols) + "values");
            out[(size_t)y * cols + x] = parse_value<T>(line.substr(start, end - start));
            ++x;
            start = end + 1;
            if (end == line.size()) break;
        }
        if (x != cols)
            throw std::runtime_error("Line" + std::to_string(y) + " is too short, expected " + std::to_string(cols) +
                                     ", got " + std::to_string(x) + " values");
    }
}

template <typename T>
matrix_array<T> load_csv(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Could not open file " + path);
    uint32_t rows, cols, count;
    read_header(in, rows, cols, count);
    matrix_array<T> a(rows, cols, count);
    for (uint32_t i = 0; i < count; ++i) read_data(in, rows, cols, a.matrix(i));
    return a;
}

// store_to_csv, src/matrix.hpp:505-511, with the row index fixed for non-square results
// (the reference indexes data[y*rows + x]) and values printed with enough digits to
// round-trip (the reference prints 6 fixed decimals).
template <typename T>
void store_csv(std::ostream& out, const matrix_array<T>& a) {
    out << "# " << a.rows << "," << a.cols << "," << a.count << "\n";
    char buf[40];
    const char* fmt = sizeof(T) == 4 ? "%.9g" : "%.17g";
    for (uint32_t i = 0; i < a.count; ++i) {
        const T* m = a.matrix(i);
        for (uint32_t y = 0; y < a.rows; ++y) {
            for (uint32_t x = 0; x < a.cols; ++x) {
                std::snprintf(buf, sizeof(buf), fmt, (double)m[(size_t)y * a.cols + x]);
                if (x) out << ',';
                out << buf;
            }
            out << '\n';
        }
    }
}

template <typename T>
void store_csv(const std::string& path, const matrix_array<T>& a) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("Could not open output file " + path);
    store_csv(out, a);
}

// get_interleaved_index, src/matrix.hpp:614-618.
inline uint32_t interleaved_index(uint32_t idx, uint32_t group_size, uint32_t num_groups) {
    return (idx % group_size) * num_groups + idx / group_size;
}

// normalize_fft_results, src/fft_helpers.hpp:193-261 (host side, for stored and validated
// FFT outputs): quadrant swap around the empty row/column, divide by rows*cols.
template <typename T>
matrix_array<T> normalize_fft(const matrix_array<T>& res) {
    matrix_array<T> norm(res.rows - 1, res.cols - 1, res.count);
    const uint32_t nr = norm.rows, nc = norm.cols, ey = res.rows / 2, ex = res.cols / 2;
    const double area = (double)res.rows * (double)res.cols;
    for (uint32_t b = 0; b < res.count; ++b) {
        const T* src = res.matrix(b);
        T* dst = norm.matrix(b);
        for (uint32_t ny = 0; ny < nr; ++ny) {
            const uint32_t sy = ny >= nr / 2 ? ny - nr / 2 : ny + ey + 1;
            for (uint32_t nx = 0; nx < nc; ++nx) {
                const uint32_t sx = nx >= nc / 2 ? nx - nc / 2 : nx + ex + 1;
                dst[(size_t)ny * nc + nx] = (T)(src[(size_t)sy * res.cols + sx] / area);
            }
        }
    }
    return norm;
}

}  // namespace xcli
