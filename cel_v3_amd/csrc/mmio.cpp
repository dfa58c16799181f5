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

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/cel_gpu.h"

namespace cel {
int fail(int code, const std::string& msg);
}

namespace {

using cel::fail;

// The file mapped into memory (read whole when it cannot be mapped); records
// are lines, addressed by pointers, so parsing allocates nothing per entry.
struct MMFile {
    const char* data = nullptr;
    size_t size = 0;
    bool mapped = false;
    std::vector<char> copy;
    const char* next = nullptr;  // start of the next record
    const char* lb = nullptr;    // current record [lb, le), '\n' and a trailing '\r' removed
    const char* le = nullptr;
    const char* cur = nullptr;   // list-directed cursor within the record
    long lineno = 0;

    ~MMFile()
    {
        if (mapped) munmap(const_cast<char*>(data), size);
    }

    int open(const char* path)
    {
        int fd = ::open(path, O_RDONLY);
        if (fd < 0) return -1;
        struct stat sb;
        if (fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size > 0) {
            void* m = mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m != MAP_FAILED) {
                data = static_cast<const char*>(m);
                size = (size_t)sb.st_size;
                mapped = true;
                madvise(m, size, MADV_SEQUENTIAL);
            }
        }
        if (!mapped) {  // empty or not mappable: read it whole
            char buf[1 << 16];
            ssize_t got;
            while ((got = ::read(fd, buf, sizeof buf)) > 0) copy.insert(copy.end(), buf, buf + got);
            data = copy.data();
            size = copy.size();
        }
        ::close(fd);
        next = data;
        return 0;
    }

    bool next_record()
    {
        const char* end = data + size;
        if (next >= end) return false;
        lb = next;
        const char* nl = static_cast<const char*>(std::memchr(lb, '\n', (size_t)(end - lb)));
        le = nl ? nl : end;
        next = nl ? nl + 1 : end;
        if (le > lb && le[-1] == '\r') --le;
        cur = lb;
        ++lineno;
        return true;
    }

    std::string line() const { return std::string(lb, le); }
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
// 1: token at [tb, tb+tn); 0: end of file; <0: a form this reader refuses.
int next_value(MMFile& F, const char*& tb, size_t& tn, bool& first, bool& comma)
{
    for (;;) {
        if (first) {
            if (!F.next_record()) return 0;
            first = false;
        }
        const char*& i = F.cur;
        const char* e = F.le;
        while (i < e && is_sep(*i)) ++i;
        if (i >= e) {
            if (!F.next_record()) return 0;
            continue;
        }
        if (*i == ',') {
            if (comma) return -2;
            comma = true;
            ++i;
            continue;
        }
        const char* b = i;
        while (i < e && !is_sep(*i) && *i != ',') ++i;
        tb = b;
        tn = (size_t)(i - b);
        for (const char* c = b; c < i; ++c) {
            if (*c == '/') return -3;
            if (*c == '*') return -4;
        }
        // a comma right after the value is its separator
        const char* j = i;
        while (j < e && is_sep(*j)) ++j;
        comma = j < e && *j == ',';
        if (comma) i = j + 1;
        return 1;
    }
}

bool parse_int(const char* t, size_t n, long long& v)
{
    size_t i = 0;
    bool neg = false;
    if (i < n && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    if (i >= n) return false;
    long long x = 0;
    for (; i < n; ++i) {
        if (t[i] < '0' || t[i] > '9') return false;
        x = x * 10 + (t[i] - '0');
        if (x > (long long)INT_MAX + 1) return false;
    }
    v = neg ? -x : x;
    return v >= INT_MIN && v <= INT_MAX;
}

// Fortran real input: [s] digits [. digits] [exp] | [s] . digits [exp],
// exp = (E|D|Q)[s]digits | s digits; also Inf/Infinity/NaN.  Rewritten to
// C form in a local buffer and converted by strtod (correct rounding).
bool parse_real(const char* t, size_t n, double& v)
{
    char small[96];
    std::string big;
    char* c = small;
    if (n + 2 > sizeof small) {
        big.resize(n + 2);
        c = &big[0];
    }
    size_t i = 0, o = 0;
    auto low = [&](size_t k) { char ch = t[k]; return (ch >= 'A' && ch <= 'Z') ? (char)(ch - 'A' + 'a') : ch; };
    if (i < n && (t[i] == '+' || t[i] == '-')) c[o++] = t[i++];
    // Inf / Infinity / NaN / NaN(...)
    if (i < n && (low(i) == 'i' || low(i) == 'n')) {
        std::string rest;
        for (size_t k = i; k < n; ++k) rest.push_back(low(k));
        bool ok = rest == "inf" || rest == "infinity" || rest == "nan" ||
                  (rest.size() > 4 && rest.compare(0, 4, "nan(") == 0 && rest.back() == ')');
        if (!ok) return false;
        for (char ch : rest) c[o++] = ch;
        c[o] = 0;
        v = std::strtod(c, nullptr);
        return true;
    }
    size_t digits = 0;
    while (i < n && t[i] >= '0' && t[i] <= '9') { c[o++] = t[i++]; ++digits; }
    if (i < n && t[i] == '.') {
        c[o++] = t[i++];
        while (i < n && t[i] >= '0' && t[i] <= '9') { c[o++] = t[i++]; ++digits; }
    }
    if (digits == 0) return false;
    if (i < n) {
        char l = low(i);
        if (l == 'e' || l == 'd' || l == 'q') ++i;
        else if (l != '+' && l != '-') return false;
        c[o++] = 'e';
        if (i < n && (t[i] == '+' || t[i] == '-')) c[o++] = t[i++];
        size_t ed = 0;
        while (i < n && t[i] >= '0' && t[i] <= '9') { c[o++] = t[i++]; ++ed; }
        if (ed == 0 || i != n) return false;
    }
    c[o] = 0;
    v = std::strtod(c, nullptr);
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
template <typename What>
int read_list(MMFile& F, int k, long long* iv, int m, double* rv, What what)
{
    bool first = true, comma = true;
    const char* tb = nullptr;
    size_t tn = 0;
    for (int j = 0; j < k + m; ++j) {
        int got = next_value(F, tb, tn, first, comma);
        if (got != 1) return read_error(got, F, what());
        bool ok = j < k ? parse_int(tb, tn, iv[j]) : parse_real(tb, tn, rv[j - k]);
        if (!ok) return read_error(-1, F, what() + " ('" + std::string(tb, tn) + "')");
    }
    return CEL_OK;
}

// mminfo: header, comments, size line.  Leaves F positioned after the size line.
int read_header(MMFile& F, const char* path, Header& H)
{
    if (F.open(path) != 0)
        return fail(CEL_ERR_FILE, std::string("cannot open ") + path + ": " + std::strerror(errno));
    if (!F.next_record()) return fail(CEL_ERR_FILE, "Premature end-of-file: no lines found");
    std::vector<std::string> w = header_words(F.line());
    if (w.size() < 5) return fail(CEL_ERR_FILE, "Invalid matrix header: " + F.line());
    if (fit(w[0], 15, false) != "%%MatrixMarket") return fail(CEL_ERR_FILE, "Invalid matrix header: " + F.line());
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
    } while (F.le > F.lb && F.lb[0] == '%');
    size_t words = header_words(F.line()).size();
    size_t want = H.rep == CEL_MM_COORDINATE ? 3 : 2;
    if (words != want)
        return fail(CEL_ERR_FILE, "Size info inconsistent with representation: " + std::to_string(want) +
                                      " size descriptors needed, " + std::to_string(words) + " found");
    // the size line is read list-directed from its own record
    std::string rec = F.line();
    long long iv[3] = {0, 0, 0};
    {
        MMFile S;  // the saved record alone, so F stays where it is
        S.data = rec.data();
        S.size = rec.size();
        S.next = S.data;
        S.next_record();
        const char* tb = nullptr;
        size_t tn = 0;
        bool first = false, comma = true;
        for (size_t j = 0; j < want; ++j) {
            int got = next_value(S, tb, tn, first, comma);
            if (got != 1 || !parse_int(tb, tn, iv[j]))
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
        rc = read_list(F, nint, iv, nreal, rv, [&]() {
            return "entry " + std::to_string(k + 1) + " of " + std::to_string(H.nnz);
        });
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
