/* cgbalance -- command-line front end of the balancer, with the flags and
 * output files of the reference's util/analyze (analyze.cpp:374-728):
 *
 *   cgbalance --matrix A.mtx (--symmetric | --shared-perm) --fabx P --faby P
 *             [--omatrix B.mtx] [--operm B.operm] [--remove_dup] [--rand K]
 *             [--threads T]
 *
 * Input: a Matrix Market "matrix coordinate real" file, entries taken as
 * stored (symmetric storage is not expanded), like util/mmio.c's
 * mm_read_unsymmetric_sparse (:16-75).
 * --omatrix: B = A(q, q) as "%%MatrixMarket matrix coordinate real general",
 *   the padded size on the size line, the input's entries in input order,
 *   1-based, values printed as C++ streams print a double (analyze.cpp:51-81).
 * --operm: the padded vertex count of lines, line i holding q[i], 0-based
 *   (analyze.cpp:700-715).
 * --rand K: a diagnostic: the heaviest PE under a random shared relabelling
 *   (seed K).  As in the reference (whose balancing restarts from the
 *   identity, analyze.cpp:252-255) it does not change the outputs.
 * --threads T: worker threads (default: hardware threads, at most 16).
 *
 * Exit status 0 on success, 1 on bad arguments, unreadable input or a
 * failed check (the reference's argument errors exit 1, :441-480).
 */
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "cgbalance.h"

namespace {

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "cgbalance: %s\n", msg.c_str());
  std::exit(1);
}

/* Buffered reader: whole lines for the header, whitespace-separated tokens
 * for the entries (fscanf's view of the file, which ignores line breaks). */
class Reader {
 public:
  explicit Reader(const std::string &path) : buf_(1 << 24) {
    f_ = std::fopen(path.c_str(), "rb");
    if (!f_) die("cannot open " + path + ": " + std::strerror(errno));
  }
  ~Reader() {
    if (f_) std::fclose(f_);
  }
  bool line(std::string &out) {
    out.clear();
    for (;;) {
      if (pos_ == len_ && !fill()) return !out.empty();
      const char *b = buf_.data() + pos_, *e = buf_.data() + len_;
      const char *nl = static_cast<const char *>(std::memchr(b, '\n', e - b));
      if (nl) {
        out.append(b, nl);
        pos_ = nl + 1 - buf_.data();
        if (!out.empty() && out.back() == '\r') out.pop_back();
        return true;
      }
      out.append(b, e);
      pos_ = len_;
    }
  }
  /* Next token; b/n point into the buffer until the next call. */
  bool token(const char *&b, size_t &n) {
    for (;;) {
      while (pos_ < len_ && is_space(buf_[pos_])) ++pos_;
      if (pos_ == len_) {
        if (!fill()) return false;
        continue;
      }
      size_t e = pos_;
      while (e < len_ && !is_space(buf_[e])) ++e;
      if (e == len_ && !eof_) {  // the token may continue past the buffer
        if (pos_ == 0) die("a token longer than 16 MiB");
        fill();
        continue;
      }
      b = buf_.data() + pos_;
      n = e - pos_;
      pos_ = e;
      return true;
    }
  }

 private:
  static bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }
  bool fill() {  // keeps the unread tail, appends more
    if (eof_) return false;
    std::memmove(buf_.data(), buf_.data() + pos_, len_ - pos_);
    len_ -= pos_;
    pos_ = 0;
    size_t got = std::fread(buf_.data() + len_, 1, buf_.size() - len_, f_);
    if (got == 0) eof_ = true;
    len_ += got;
    return got > 0;
  }
  std::FILE *f_ = nullptr;
  std::vector<char> buf_;
  size_t pos_ = 0, len_ = 0;
  bool eof_ = false;
};

int64_t parse_int(const char *b, size_t n, const char *what) {
  size_t i = 0;
  bool neg = false;
  if (i < n && (b[i] == '+' || b[i] == '-')) neg = b[i++] == '-';
  if (i == n) die(std::string("expected an integer for ") + what);
  int64_t v = 0;
  for (; i < n; ++i) {
    if (b[i] < '0' || b[i] > '9') die(std::string("expected an integer for ") + what);
    v = v * 10 + (b[i] - '0');
    if (v > (int64_t)1 << 40) die(std::string("integer out of range for ") + what);
  }
  return neg ? -v : v;
}

double parse_value(const char *b, size_t n) {
  char tmp[128];
  if (n >= sizeof(tmp)) die("a value token is too long");
  std::memcpy(tmp, b, n);
  tmp[n] = 0;
  char *end = nullptr;
  double v = std::strtod(tmp, &end);
  if (end != tmp + n) die(std::string("bad value '") + tmp + "'");
  return v;
}

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

struct Coo {
  int64_t m = 0, n = 0, nnz = 0;
  std::vector<int32_t> row, col;
  std::vector<double> val;
};

/* A "matrix coordinate real" file; other types are refused, as
 * mm_read_unsymmetric_sparse refuses them (mmio.c:34-39). */
Coo read_mtx(const std::string &path, bool keep_values) {
  Reader rd(path);
  std::string ln;
  if (!rd.line(ln)) die(path + ": empty file");
  char banner[64] = {0}, object[64] = {0}, format[64] = {0}, field[64] = {0}, sym[64] = {0};
  if (std::sscanf(ln.c_str(), "%63s %63s %63s %63s %63s", banner, object, format, field, sym) != 5 ||
      std::string(banner) != "%%MatrixMarket")
    die(path + ": not a Matrix Market file");
  const std::string o = lower(object), f = lower(format), d = lower(field), s = lower(sym);
  if (o != "matrix" || f != "coordinate" || d != "real" ||
      (s != "general" && s != "symmetric" && s != "skew-symmetric" && s != "hermitian"))
    die(path + ": unsupported Matrix Market type [" + o + " " + f + " " + d + " " + s +
        "]; a real coordinate matrix is required");
  do {
    if (!rd.line(ln)) die(path + ": no size line");
  } while (ln.empty() || ln[0] == '%');
  long long M = 0, N = 0, Z = 0;
  if (std::sscanf(ln.c_str(), "%lld %lld %lld", &M, &N, &Z) != 3 || M < 0 || N < 0 || Z < 0 ||
      M > INT32_MAX || N > INT32_MAX || Z > INT32_MAX)
    die(path + ": bad size line '" + ln + "'");
  Coo a;
  a.m = M;
  a.n = N;
  a.nnz = Z;
  a.row.resize(Z);
  a.col.resize(Z);
  if (keep_values) a.val.resize(Z);
  const char *b;
  size_t n;
  for (int64_t i = 0; i < Z; ++i) {
    if (!rd.token(b, n)) die(path + ": fewer entries than the size line says");
    const int64_t r = parse_int(b, n, "a row index");
    if (!rd.token(b, n)) die(path + ": truncated entry");
    const int64_t c = parse_int(b, n, "a column index");
    if (!rd.token(b, n)) die(path + ": truncated entry");
    if (keep_values) a.val[i] = parse_value(b, n);
    if (r < 1 || r > M || c < 1 || c > N)
      die(path + ": entry " + std::to_string(i + 1) + " lies outside the matrix");
    a.row[i] = (int32_t)(r - 1);
    a.col[i] = (int32_t)(c - 1);
  }
  return a;
}

class Writer {
 public:
  explicit Writer(const std::string &path) : path_(path) {
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) die("cannot write " + path + ": " + std::strerror(errno));
    buf_.reserve(kCap + 256);
  }
  ~Writer() { close(); }
  void put(const char *s, size_t n) {
    buf_.append(s, n);
    if (buf_.size() >= kCap) flush();
  }
  void put(const std::string &s) { put(s.data(), s.size()); }
  void put_int(int64_t v) {
    char t[24];
    int k = 0;
    bool neg = v < 0;
    uint64_t u = neg ? 0 - (uint64_t)v : (uint64_t)v;
    do {
      t[k++] = (char)('0' + u % 10);
      u /= 10;
    } while (u);
    if (neg) t[k++] = '-';
    std::reverse(t, t + k);
    put(t, k);
  }
  /* A double as an ostream with default flags prints it: printf's %g with
   * precision 6.  Integral values below 1e6 in magnitude print as integers
   * under %g, which is the common case here and skips snprintf. */
  void put_double(double v) {
    if (v == std::trunc(v) && std::fabs(v) < 1e6 && !(v == 0 && std::signbit(v))) {
      put_int((int64_t)v);
      return;
    }
    char t[64];
    int k = std::snprintf(t, sizeof(t), "%g", v);
    put(t, (size_t)k);
  }
  void close() {
    if (!f_) return;
    flush();
    if (std::fclose(f_) != 0) die("write to " + path_ + " failed");
    f_ = nullptr;
  }

 private:
  static constexpr size_t kCap = 1 << 22;
  void flush() {
    if (!buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
      die("write to " + path_ + " failed");
    buf_.clear();
  }
  std::string path_;
  std::FILE *f_ = nullptr;
  std::string buf_;
};

/* Heaviest PE when vertex v goes to block pos[v] / block on both axes. */
int64_t heaviest_pe(const Coo &a, const std::vector<int64_t> &pos, int64_t block, int grid) {
  std::vector<int64_t> cnt((size_t)grid * grid, 0);
  for (int64_t i = 0; i < a.nnz; ++i)
    cnt[(size_t)(pos[a.col[i]] / block) * grid + (size_t)(pos[a.row[i]] / block)]++;
  return *std::max_element(cnt.begin(), cnt.end());
}

struct Args {
  std::string matrix, omatrix, operm;
  int fabx = 0, faby = 0, rand = 0, threads = 0;
  bool remove_dup = false, symmetric = false, shared_perm = false;
};

void usage() {
  std::fprintf(stderr,
               "usage: cgbalance --matrix A.mtx (--symmetric | --shared-perm) --fabx P --faby P\n"
               "                 [--omatrix B.mtx] [--operm B.operm] [--remove_dup] [--rand K]\n"
               "                 [--threads T]\n");
}

Args parse_args(int argc, char **argv) {
  Args a;
  auto value = [&](int &i) -> std::string {
    if (i + 1 >= argc) {
      usage();
      die(std::string(argv[i]) + " needs a value");
    }
    return argv[++i];
  };
  auto number = [&](int &i) -> int {
    std::string s = value(i);
    char *end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (s.empty() || *end || v < INT32_MIN || v > INT32_MAX) die("bad integer '" + s + "'");
    return (int)v;
  };
  for (int i = 1; i < argc; ++i) {
    std::string k = argv[i];
    if (k == "--matrix") a.matrix = value(i);
    else if (k == "--omatrix") a.omatrix = value(i);
    else if (k == "--operm") a.operm = value(i);
    else if (k == "--fabx") a.fabx = number(i);
    else if (k == "--faby") a.faby = number(i);
    else if (k == "--rand") a.rand = number(i);
    else if (k == "--threads") a.threads = number(i);
    else if (k == "--remove_dup") a.remove_dup = true;
    else if (k == "--symmetric") a.symmetric = true;
    else if (k == "--shared-perm") a.shared_perm = true;
    else if (k == "-h" || k == "--help") {
      usage();
      std::exit(0);
    } else {
      usage();
      die("unknown argument " + k);
    }
  }
  if (a.matrix.empty()) {
    usage();
    die("--matrix is required");
  }
  if (a.symmetric && a.shared_perm)
    die("--symmetric and --shared-perm are mutually exclusive (--symmetric is --shared-perm plus "
        "the structural-symmetry check)");
  if (!a.symmetric && !a.shared_perm)
    die("one of --symmetric or --shared-perm is required: the permutation is always shared by "
        "rows and columns, so say whether the input is structurally symmetric");
  if (a.fabx <= 0 || a.faby <= 0) die("--fabx and --faby must be positive");
  return a;
}

}  // namespace

int main(int argc, char **argv) {
  const Args args = parse_args(argc, argv);
  std::printf("reading %s\n", args.matrix.c_str());
  std::fflush(stdout);
  Coo a = read_mtx(args.matrix, !args.omatrix.empty());
  if (a.m != a.n || args.fabx != args.faby)
    die("--symmetric/--shared-perm need a square matrix and a square PE grid: got " +
        std::to_string(a.m) + "x" + std::to_string(a.n) + " matrix, " + std::to_string(args.fabx) +
        "x" + std::to_string(args.faby) + " grid");
  if (a.n == 0) die("the matrix has no rows");
  const int64_t npad = cgbal_padded_n(a.n, args.fabx, args.faby);
  if (npad < 0) die(cgbal_last_error());
  std::printf("n = %lld, nnz = %lld, grid %d x %d, padded n = %lld\n", (long long)a.n,
              (long long)a.nnz, args.fabx, args.faby, (long long)npad);

  std::vector<int64_t> q(npad);
  const int32_t flags = (args.symmetric ? CGBAL_SYMMETRIC : CGBAL_SHARED_PERM) |
                        (args.remove_dup ? CGBAL_REMOVE_DUP : 0);
  cgbal_report rep;
  std::vector<int64_t> buckets((size_t)std::min<int64_t>(args.fabx, a.n) * std::min<int64_t>(args.fabx, a.n));
  int rc = cgbal_balance_coo(a.n, a.nnz, a.row.data(), a.col.data(), args.fabx, args.faby, flags,
                             args.threads, q.data(), npad, buckets.data(), &rep);
  if (rc != CGBAL_OK) die(cgbal_last_error());

  if (args.rand) {
    std::vector<int64_t> pos(npad);
    std::iota(pos.begin(), pos.end(), 0);
    std::mt19937_64 gen((uint64_t)args.rand);
    std::shuffle(pos.begin(), pos.end(), gen);
    std::printf("random relabelling (seed %d): heaviest PE has %lld nz (diagnostic only)\n",
                args.rand, (long long)heaviest_pe(a, pos, rep.block, rep.grid));
  }
  std::printf("block %lld x %lld%s, %lld entries balanced\n", (long long)rep.block,
              (long long)rep.block, rep.exact_blocks ? " (exact ceiling)" : "", (long long)rep.nnz);
  std::printf("identity placement: heaviest PE has %lld nz\n", (long long)rep.max_identity);
  std::printf("balanced placement: heaviest PE has %lld nz, balance %.6f, %d passes\n",
              (long long)rep.max_balanced, rep.lb, rep.passes);

  if (!args.omatrix.empty()) {
    Writer w(args.omatrix);
    w.put("%%MatrixMarket matrix coordinate real general\n");
    w.put_int(npad);
    w.put(" ", 1);
    w.put_int(npad);
    w.put(" ", 1);
    w.put_int(a.nnz);
    w.put("\n", 1);
    for (int64_t i = 0; i < a.nnz; ++i) {
      w.put_int(q[a.row[i]] + 1);
      w.put(" ", 1);
      w.put_int(q[a.col[i]] + 1);
      w.put(" ", 1);
      w.put_double(a.val[i]);
      w.put("\n", 1);
    }
    w.close();
    std::printf("wrote %s\n", args.omatrix.c_str());
  }
  if (!args.operm.empty()) {
    Writer w(args.operm);
    for (int64_t i = 0; i < npad; ++i) {
      w.put_int(q[i]);
      w.put("\n", 1);
    }
    w.close();
    std::printf("wrote %s\n", args.operm.c_str());
  }
  return 0;
}
