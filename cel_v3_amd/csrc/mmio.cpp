// mmio.cpp -- Matrix Market reader behind cel_gpu_mm_info / cel_gpu_mm_read.
//
// Host code only (parsing a text file is not GPU work): it fills host COO
// arrays that cel_gpu_coo_sort and cel_gpu_mat_create_csr then move to HBM.
//
// It accepts and rejects what the reference's reader accepts and rejects
// and returns the same entries:
//   mminfo  src/fbase/cel_mmio_module.f90:342-466
//   mmread  src/fbase/cel_mmio_module.f90:17-339
//   getwd / countwd / lowerc  :766-845 (header words)
//   the caller  cel_sp_mat_read_matrix_market_file
//               src/fspmt/cel_sp_mat_generate_module.f90:219-273
//
// Header line: words are maximal runs of characters other than ' ' (a tab is
// part of a word) within the first 1024 characters; each word is cut to the
// width of the reference's variable (15, 6, 10, 7, 19) and compared with
// blanks padded, after lowering the case of all but the first.  So
// "coordinates" reads as "coordinate" and "reals" is not "real".  The field
// 'integer' is never accepted: the reference compares its 7-character field
// with 'integer(kind=ik)'.
// Comment lines start with '%' in column 1; the first other line is the size
// line, which must hold exactly 3 (coordinate) or 2 (array) blank-separated
// words.  Values are read as Fortran list-directed input, one READ per
// entry: separators are blanks, tabs or one comma; a READ continues onto the
// next lines until it has its values and drops the rest of its last line; a
// real may use E, D or Q exponent letters or a signed exponent without a
// letter.  The list-directed forms r*c (repeat), null values and '/' are
// refused with CEL_ERR_FILE rather than interpreted.  A trailing '\r' is
// dropped from every line.
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cel_gpu.h"

namespace cel {
int fail(int code, const std::string& msg);
}

namespace {

using cel::fail;

struct MMFile {
    FILE* f = nullptr;
    std::string line;  // current record, '\n' and a trailing '\r' removed
    size_t pos = 0;    // list-directed cursor within line
    long lineno = 0;

    ~MMFile() { if (f) std::fclose(f); }

    bool next_record()
    {
        line.clear();
        pos = 0;
        int c;
        bool any = false;
        while ((c = std::fgetc(f)) != EOF) {
            any = true;
            if (c == '\n') break;
            line.push_back((char)c);
        }
        if (!any) return false;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        ++lineno;
        return true;
    }
};

// Word j (0-based) of the first 1024 characters, split at ' ' only.
std::vector<std::string> header_words(const std::string& line)
{
    std::string s = line.substr(0, 1024);
    std::vector<std::string> w;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && s[i] == ' ') ++i;
        size_t b = i;
        while (i < s.size() && s[i] != ' ') ++i;
        if (i > b) w.push_back(s.substr(b, i - b));
    }
    return w;
}

std::string fit(const std::string& w, size_t width, bool lower)
{
    std::string t = w.substr(0, width);
    if (lower)
        for (char& c : t)
            if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    return t;
}

struct Header {
    int rep = 0, field = 0, symm = 0;
    long long nrows = 0, ncols = 0, nnz = 0;
};

bool is_sep(char c) { return c == ' ' || c == '\t'; }

// Next list-directed value of the current READ; crosses records as needed.
// `first` starts a new record (a READ begins with one); `comma` tells that
// the previous value was closed by a comma, so a second comma (or a comma
// opening the READ) marks a null value.
// 1: token in tok; 0: end of file; <0: a form this reader refuses.
int next_value(MMFile& F, std::string& tok, bool& first, bool& comma)
{
    for (;;) {
        if (first) {
            if (!F.next_record()) return 0;
            first = false;
        }
        std::string& s = F.line;
        size_t& i = F.pos;
        while (i < s.size() && is_sep(s[i])) ++i;
        if (i >= s.size()) {
            if (!F.next_record()) return 0;
            continue;
        }
        if (s[i] == ',') {
            if (comma) return -2;
            comma = true;
            ++i;
            continue;
        }
        size_t b = i;
        while (i < s.size() && !is_sep(s[i]) && s[i] != ',') ++i;
        tok = s.substr(b, i - b);
        if (tok.find('/') != std::string::npos) return -3;
        if (tok.find('*') != std::string::npos) return -4;
        // a comma right after the value is its separator
        size_t j = i;
        while (j < s.size() && is_sep(s[j])) ++j;
        comma = j < s.size() && s[j] == ',';
        if (comma) i = j + 1;
        return 1;
    }
}

bool parse_int(const std::string& t, long long& v)
{
    size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    if (i >= t.size()) return false;
    long long x = 0;
    for (; i < t.size(); ++i) {
        if (t[i] < '0' || t[i] > '9') return false;
        x = x * 10 + (t[i] - '0');
        if (x > (long long)INT_MAX + 1) return false;
    }
    v = neg ? -x : x;
    return v >= INT_MIN && v <= INT_MAX;
}

// Fortran real input: [s] digits [. digits] [exp] | [s] . digits [exp],
// exp = (E|D|Q)[s]digits | s digits; also Inf/Infinity/NaN.  Converted by
// strtod (correct rounding) after mapping the exponent to C form.
bool parse_real(const std::string& t, double& v)
{
    std::string low;
    for (char c : t) low.push_back((c >= 'A' && c <= 'Z') ? (char)(c - 'A' + 'a') : c);
    size_t i = 0;
    std::string c;
    if (i < low.size() && (low[i] == '+' || low[i] == '-')) c.push_back(low[i++]);
    std::string rest = low.substr(i);
    if (rest == "inf" || rest == "infinity" || rest == "nan" ||
        (rest.size() > 4 && rest.compare(0, 4, "nan(") == 0 && rest.back() == ')')) {
        v = std::strtod((c + rest).c_str(), nullptr);
        return true;
    }
    size_t digits = 0;
    while (i < low.size() && low[i] >= '0' && low[i] <= '9') { c.push_back(low[i++]); ++digits; }
    if (i < low.size() && low[i] == '.') {
        c.push_back(low[i++]);
        while (i < low.size() && low[i] >= '0' && low[i] <= '9') { c.push_back(low[i++]); ++digits; }
    }
    if (digits == 0) return false;
    if (i < low.size()) {
        if (low[i] == 'e' || low[i] == 'd' || low[i] == 'q') ++i;
        else if (low[i] != '+' && low[i] != '-') return false;
        c.push_back('e');
        if (i < low.size() && (low[i] == '+' || low[i] == '-')) c.push_back(low[i++]);
        size_t ed = 0;
        while (i < low.size() && low[i] >= '0' && low[i] <= '9') { c.push_back(low[i++]); ++ed; }
        if (ed == 0 || i != low.size()) return false;
    }
    v = std::strtod(c.c_str(), nullptr);
    return true;
}

int read_error(int got, const MMFile& F, const std::string& what)
{
    if (got == 0)
        return fail(CEL_ERR_FILE, "Premature end-of-file while reading " + what);
    const char* why = got == -2 ? "a null value" : got == -3 ? "a '/' terminator"
                    : got == -4 ? "a repeat count (r*c)" : "an unreadable value";
    return fail(CEL_ERR_FILE, std::string("line ") + std::to_string(F.lineno) + ": " + why +
                                  " while reading " + what);
}

// One list-directed READ of k integers and m reals.
int read_list(MMFile& F, int k, long long* iv, int m, double* rv, const std::string& what)
{
    bool first = true, comma = true;
    std::string tok;
    for (int j = 0; j < k + m; ++j) {
        int got = next_value(F, tok, first, comma);
        if (got != 1) return read_error(got, F, what);
        bool ok = j < k ? parse_int(tok, iv[j]) : parse_real(tok, rv[j - k]);
        if (!ok) return read_error(-1, F, what + " ('" + tok + "')");
    }
    return CEL_OK;
}

// mminfo: header, comments, size line.  Leaves F positioned after the size line.
int read_header(MMFile& F, const char* path, Header& H)
{
    F.f = std::fopen(path, "r");
    if (!F.f) return fail(CEL_ERR_FILE, std::string("cannot open ") + path + ": " + std::strerror(errno));
    if (!F.next_record()) return fail(CEL_ERR_FILE, "Premature end-of-file: no lines found");
    std::vector<std::string> w = header_words(F.line);
    if (w.size() < 5) return fail(CEL_ERR_FILE, "Invalid matrix header: " + F.line);
    if (fit(w[0], 15, false) != "%%MatrixMarket") return fail(CEL_ERR_FILE, "Invalid matrix header: " + F.line);
    if (fit(w[1], 6, true) != "matrix") return fail(CEL_ERR_FILE, "Invalid matrix type: " + w[1]);
    std::string rep = fit(w[2], 10, true), field = fit(w[3], 7, true), symm = fit(w[4], 19, true);
    if (rep == "coordinate") H.rep = CEL_MM_COORDINATE;
    else if (rep == "array") H.rep = CEL_MM_ARRAY;
    else return fail(CEL_ERR_FILE, "'" + rep + "' representation not recognized");
    if (field == "real") H.field = CEL_MM_REAL;
    else if (field == "complex") H.field = CEL_MM_COMPLEX;
    else if (field == "pattern" && H.rep == CEL_MM_COORDINATE) H.field = CEL_MM_PATTERN;
    else return fail(CEL_ERR_FILE, "'" + field + "' field is not recognized" +
                                       std::string(field == "integer" ? " (the reference compares it "
                                                   "with 'integer(kind=ik)' and never matches)" : ""));
    if (symm == "general") H.symm = CEL_MM_GENERAL;
    else if (symm == "symmetric") H.symm = CEL_MM_SYMMETRIC;
    else if (symm == "hermitian") H.symm = CEL_MM_HERMITIAN;
    else if (symm == "skew-symmetric") H.symm = CEL_MM_SKEW_SYMMETRIC;
    else return fail(CEL_ERR_FILE, "'" + symm + "' symmetry is not recognized");
    // comments: '%' in column 1; the first other line is the size line
    do {
        if (!F.next_record()) return fail(CEL_ERR_FILE, "Premature end-of-file: no data lines found");
    } while (!F.line.empty() && F.line[0] == '%');
    size_t words = header_words(F.line).size();
    size_t want = H.rep == CEL_MM_COORDINATE ? 3 : 2;
    if (words != want)
        return fail(CEL_ERR_FILE, "Size info inconsistent with representation: " + std::to_string(want) +
                                      " size descriptors needed, " + std::to_string(words) + " found");
    // the size line is read list-directed from its own record
    std::string rec = F.line;
    long long iv[3] = {0, 0, 0};
    {
        MMFile S;  // read from the saved record without advancing F
        S.line = rec;
        std::string tok;
        bool first = false, comma = true;
        for (size_t j = 0; j < want; ++j) {
            int got = next_value(S, tok, first, comma);
            if (got != 1 || !parse_int(tok, iv[j]))
                return fail(CEL_ERR_FILE, "unreadable size line: " + rec);
        }
    }
    H.nrows = iv[0];
    H.ncols = iv[1];
    if (H.nrows < 0 || H.ncols < 0) return fail(CEL_ERR_FILE, "negative matrix size: " + rec);
    if (H.rep == CEL_MM_COORDINATE) {
        H.nnz = iv[2];
        if (H.nnz < 0) return fail(CEL_ERR_FILE, "negative entry count: " + rec);
    } else {
        long long r = H.nrows, c = H.ncols;
        H.nnz = H.symm == CEL_MM_SYMMETRIC || H.symm == CEL_MM_HERMITIAN ? (r * c - r) / 2 + r
              : H.symm == CEL_MM_SKEW_SYMMETRIC ? (r * c - r) / 2 : r * c;
        if (H.nnz > INT_MAX) return fail(CEL_ERR_FILE, "array matrix too large for 32-bit counts");
    }
    return CEL_OK;
}

}  // namespace

extern "C" int cel_gpu_mm_info(const char* path, int* rep, int* field, int* symm, int* nrows,
                               int* ncols, int* nnz)
{
    if (!path) return fail(CEL_ERR_ARG, "cel_gpu_mm_info: null path");
    MMFile F;
    Header H;
    int rc = read_header(F, path, H);
    if (rc < 0) return rc;
    if (rep) *rep = H.rep;
    if (field) *field = H.field;
    if (symm) *symm = H.symm;
    if (nrows) *nrows = (int)H.nrows;
    if (ncols) *ncols = (int)H.ncols;
    if (nnz) *nnz = (int)H.nnz;
    return CEL_OK;
}

extern "C" int cel_gpu_mm_read(const char* path, int flags, int cap, int* rows, int* cols,
                               double* vals, int* nnz_out)
{
    if (!path || cap < 0 || (flags & ~CEL_MM_EXPAND_SYMMETRIC))
        return fail(CEL_ERR_ARG, "cel_gpu_mm_read: bad arguments");
    MMFile F;
    Header H;
    int rc = read_header(F, path, H);
    if (rc < 0) return rc;
    if (H.rep == CEL_MM_ARRAY)
        return fail(CEL_ERR_REF_UNDEFINED, "array (dense) file: the reference's reader fills only "
                                           "the values and leaves the row and column indices unset");
    if (H.field == CEL_MM_COMPLEX && H.nnz > 1)
        return fail(CEL_ERR_REF_UNDEFINED, "complex file: the reference reads the values into a "
                                           "one-element buffer");
    bool expand = (flags & CEL_MM_EXPAND_SYMMETRIC) && H.symm != CEL_MM_GENERAL;
    if (H.nnz > cap) return fail(CEL_ERR_ARG, "insufficient array length for " + std::to_string(H.nnz) + " entries");
    if (H.nnz > 0 && (!rows || !cols || (H.field == CEL_MM_REAL && !vals)))
        return fail(CEL_ERR_ARG, "cel_gpu_mm_read: null output array");
    const int nint = 2;
    const int nreal = H.field == CEL_MM_REAL ? 1 : H.field == CEL_MM_COMPLEX ? 2 : 0;
    for (long long k = 0; k < H.nnz; ++k) {
        long long iv[2];
        double rv[2];
        rc = read_list(F, nint, iv, nreal, rv, "entry " + std::to_string(k + 1) + " of " +
                                                  std::to_string(H.nnz));
        if (rc < 0) return rc;
        rows[k] = (int)iv[0];
        cols[k] = (int)iv[1];
        if (H.field == CEL_MM_REAL) vals[k] = rv[0];
    }
    long long total = H.nnz;
    if (expand) {
        long long extra = 0;
        for (long long k = 0; k < H.nnz; ++k) extra += rows[k] != cols[k];
        if (H.nnz + extra > cap)
            return fail(CEL_ERR_ARG, "insufficient array length for " + std::to_string(H.nnz + extra) +
                                         " entries after symmetric expansion");
        double sign = H.symm == CEL_MM_SKEW_SYMMETRIC ? -1.0 : 1.0;
        for (long long k = 0; k < H.nnz; ++k) {
            if (rows[k] == cols[k]) continue;
            rows[total] = cols[k];
            cols[total] = rows[k];
            if (H.field == CEL_MM_REAL) vals[total] = sign * vals[k];
            ++total;
        }
    }
    if (nnz_out) *nnz_out = (int)total;
    return CEL_OK;
}
