// graph_io.hip -- AdjacencyGraph files on the host, so a graph this library
// builds can be read by a reference run and the other way round.
//
// Text (parse_unweighted_graph, gbbs/graph_io.cc:35-100; written by
// write_graph_to_file, gbbs/graph_io.h:634-687): whitespace-separated
// tokens "AdjacencyGraph", n, m, the n offsets, the m neighbour ids;
// offsets[n] = m is implied.
// Binary (the -b layout read by graph_io.cc:86-97 with -DLONG): three
// int64 {n, m, bytes of the edge section}, (n+1) uint64 offsets, m uint32
// ids.
//
// Host code only: no device is needed to read or write a file.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace byo {
namespace {

constexpr const char *kHeader = "AdjacencyGraph";

struct File {
  FILE *f = nullptr;
  File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

std::string read_all(const char *path) {
  File in(path, "rb");
  BYO_REQUIRE(in.f, BYO_E_INVAL, std::string("cannot open ") + path);
  std::string s;
  char buf[1 << 16];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof(buf), in.f)) > 0) s.append(buf, r);
  BYO_REQUIRE(!std::ferror(in.f), BYO_E_INVAL, std::string("read error on ") + path);
  return s;
}

// Tokenizer over the text format: next unsigned decimal, or the header.
struct Tokens {
  const char *p, *end;
  void skip() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
  }
  bool done() {
    skip();
    return p >= end;
  }
  std::string word() {
    skip();
    const char *b = p;
    while (p < end && !(*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
    return std::string(b, p);
  }
  uint64_t number(const char *what) {
    skip();
    BYO_REQUIRE(p < end, BYO_E_INVAL, std::string("file ends before all ") + what);
    uint64_t v = 0;
    const char *b = p;
    while (p < end && *p >= '0' && *p <= '9') {
      const uint64_t d = (uint64_t)(*p - '0');
      BYO_REQUIRE(v <= (UINT64_MAX - d) / 10, BYO_E_INVAL, std::string("number too large in ") + what);
      v = v * 10 + d;
      p++;
    }
    BYO_REQUIRE(p > b && (p == end || *p == ' ' || *p == '\n' || *p == '\r' || *p == '\t'), BYO_E_INVAL,
                std::string("not an unsigned integer in ") + what);
    return v;
  }
};

void text_header(Tokens &t, uint64_t *n, uint64_t *m) {
  BYO_REQUIRE(t.word() == kHeader, BYO_E_INVAL, "not an unweighted AdjacencyGraph file");
  *n = t.number("the header");
  *m = t.number("the header");
  BYO_REQUIRE(*n < 0xFFFFFFFFull, BYO_E_INVAL, "n must be below 2^32-1 (uint32 vertex ids)");
}

struct BinHeader {
  int64_t n, m, space;
};

BinHeader bin_header(FILE *f, const char *path) {
  BinHeader h{};
  BYO_REQUIRE(std::fread(&h, sizeof(h), 1, f) == 1, BYO_E_INVAL, std::string("short binary header in ") + path);
  BYO_REQUIRE(h.n >= 0 && h.m >= 0 && (uint64_t)h.n < 0xFFFFFFFFull, BYO_E_INVAL, "bad n or m in binary header");
  return h;
}

// Writes v in decimal followed by '\n' at out; returns the new end.
char *put_line(char *out, uint64_t v) {
  char tmp[24];
  int k = 0;
  do {
    tmp[k++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (k) *out++ = tmp[--k];
  *out++ = '\n';
  return out;
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_adjacency_info(const char *path, int binary, uint64_t *n, uint64_t *m) {
  return guard([&] {
    BYO_REQUIRE(path && n && m, BYO_E_INVAL, "null argument");
    if (binary) {
      File in(path, "rb");
      BYO_REQUIRE(in.f, BYO_E_INVAL, std::string("cannot open ") + path);
      const BinHeader h = bin_header(in.f, path);
      *n = (uint64_t)h.n;
      *m = (uint64_t)h.m;
    } else {
      // the header is the first three tokens; read just enough of the file
      File in(path, "rb");
      BYO_REQUIRE(in.f, BYO_E_INVAL, std::string("cannot open ") + path);
      char buf[256];
      const size_t r = std::fread(buf, 1, sizeof(buf), in.f);
      Tokens t{buf, buf + r};
      uint64_t nn, mm;
      text_header(t, &nn, &mm);
      BYO_REQUIRE(t.p < t.end || r < sizeof(buf), BYO_E_INVAL, "header too long");
      *n = nn;
      *m = mm;
    }
  });
}

int byo_adjacency_read(const char *path, int binary, uint64_t n, uint64_t m, uint64_t *offsets, uint32_t *ids) {
  return guard([&] {
    BYO_REQUIRE(path && offsets && (ids || m == 0), BYO_E_INVAL, "null argument");
    if (binary) {
      File in(path, "rb");
      BYO_REQUIRE(in.f, BYO_E_INVAL, std::string("cannot open ") + path);
      const BinHeader h = bin_header(in.f, path);
      BYO_REQUIRE((uint64_t)h.n == n && (uint64_t)h.m == m, BYO_E_INVAL, "n or m differ from the file's header");
      BYO_REQUIRE(std::fread(offsets, 8, n + 1, in.f) == n + 1, BYO_E_INVAL, "binary file ends inside the offsets");
      BYO_REQUIRE(m == 0 || std::fread(ids, 4, m, in.f) == m, BYO_E_INVAL, "binary file ends inside the ids");
    } else {
      const std::string s = read_all(path);
      Tokens t{s.data(), s.data() + s.size()};
      uint64_t nn, mm;
      text_header(t, &nn, &mm);
      BYO_REQUIRE(nn == n && mm == m, BYO_E_INVAL, "n or m differ from the file's header");
      for (uint64_t v = 0; v < n; v++) offsets[v] = t.number("the offsets");
      offsets[n] = m;
      for (uint64_t e = 0; e < m; e++) {
        const uint64_t x = t.number("the neighbour ids");
        BYO_REQUIRE(x < n, BYO_E_INVAL, "neighbour id out of range");
        ids[e] = (uint32_t)x;
      }
      BYO_REQUIRE(t.done(), BYO_E_INVAL, "tokens after the last neighbour id");
    }
    check_csr_input(offsets, ids, n, m);
  });
}

int byo_adjacency_write(const char *path, int binary, const uint64_t *offsets, const uint32_t *ids, uint64_t n,
                        uint64_t m) {
  return guard([&] {
    BYO_REQUIRE(path, BYO_E_INVAL, "null argument");
    check_csr_input(offsets, ids, n, m);
    File out(path, "wb");
    BYO_REQUIRE(out.f, BYO_E_INVAL, std::string("cannot create ") + path);
    bool ok = true;
    if (binary) {
      const BinHeader h{(int64_t)n, (int64_t)m, (int64_t)(m * sizeof(uint32_t))};
      ok = std::fwrite(&h, sizeof(h), 1, out.f) == 1 && std::fwrite(offsets, 8, n + 1, out.f) == n + 1 &&
           (m == 0 || std::fwrite(ids, 4, m, out.f) == m);
    } else {
      // one number per line, as write_graph_to_file does, through a buffer
      std::vector<char> buf(1 << 20);
      char *p = buf.data();
      char *const lim = buf.data() + buf.size() - 64;
      auto flush = [&] {
        ok = ok && std::fwrite(buf.data(), 1, (size_t)(p - buf.data()), out.f) == (size_t)(p - buf.data());
        p = buf.data();
      };
      p += std::snprintf(p, 64, "%s\n", kHeader);
      p = put_line(p, n);
      p = put_line(p, m);
      for (uint64_t v = 0; v < n; v++) {
        if (p > lim) flush();
        p = put_line(p, offsets[v]);
      }
      for (uint64_t e = 0; e < m; e++) {
        if (p > lim) flush();
        p = put_line(p, ids[e]);
      }
      flush();
    }
    BYO_REQUIRE(ok && std::fflush(out.f) == 0, BYO_E_INVAL, std::string("write error on ") + path);
  });
}

}  // extern "C"
