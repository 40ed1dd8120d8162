// Reader for MATLAB -v7.3 MAT-files: HDF5 files behind a 512-byte MATLAB user
// block.  MesoLF's PSF files are such files; mesolf.m:20-21 loads one whole
// with load(SI.psffile) and reads H, Ht, Nnum (recon:10-16 consume H, Ht),
// CAindex (forwardProjectACC.m:12) and x1objspace / x3objspace (mesolf.m:231-232).
//
// The subset decoded is what MATLAB and the HDF5 library's default settings
// write for real numeric and logical arrays:
//   superblock versions 0-3 (addresses relative to the superblock's base),
//   object headers version 1 and 2 with continuation blocks,
//   groups as symbol tables (v1 B-tree + local heap) or compact link messages,
//   dataspaces v1/v2, fixed-point and IEEE float datatypes of either byte
//   order, string attributes (MATLAB_class, MATLAB_empty),
//   layouts v3/v4: compact, contiguous, chunked with a v1 B-tree, single-chunk,
//   implicit or unpaged fixed-array index,
//   filters deflate, shuffle and fletcher32, fill values.
// Complex (compound), sparse, cell, struct, object and dense-link-storage
// variables are refused with MLF_ERR_UNSUPPORTED and a message naming them.
// Every read is bounds-checked against the file; a malformed file gives
// MLF_ERR_IO, never an out-of-bounds access.
//
// Output order: MATLAB's dims are the HDF5 dims reversed and its column-major
// data is the HDF5 row-major data, byte for byte, so a variable is returned in
// MATLAB order without a transpose.  Chunks decode in parallel.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace mlf {
namespace mat73 {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  mlf_status code;
  Error(mlf_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] static void fail(mlf_status code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

constexpr uint64_t kUndef = ~0ull;                 // HDF5 "undefined address"
constexpr size_t kMaxMeta = 64ull << 20;           // largest metadata block read
constexpr int kMaxDepth = 32;                      // B-tree / group nesting bound

// ---------------------------------------------------------------- cursor
struct Cur {
  const uint8_t* p;
  const uint8_t* e;
  const char* what;
  Cur(const uint8_t* b, size_t n, const char* w) : p(b), e(b + n), what(w) {}
  size_t left() const { return (size_t)(e - p); }
  void need(size_t n) const {
    if (left() < n) fail(MLF_ERR_IO, "truncated %s", what);
  }
  uint64_t un(int n) {
    need((size_t)n);
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v |= (uint64_t)p[i] << (8 * i);
    p += n;
    return v;
  }
  uint8_t u8() { return (uint8_t)un(1); }
  uint16_t u16() { return (uint16_t)un(2); }
  uint32_t u32() { return (uint32_t)un(4); }
  uint64_t u64() { return un(8); }
  void skip(size_t n) {
    need(n);
    p += n;
  }
  const uint8_t* take(size_t n) {
    need(n);
    const uint8_t* q = p;
    p += n;
    return q;
  }
};

// --------------------------------------------------------------- the file
struct File {
  int fd = -1;
  uint64_t size = 0;
  uint64_t base = 0;  // absolute offset of the superblock (addresses are relative to it)
  int so = 8, sl = 8; // size of offsets, size of lengths
  uint64_t root = kUndef;  // root group object header (relative address)
  std::string path;

  ~File() {
    if (fd >= 0) ::close(fd);
  }

  void pread_abs(uint64_t off, void* dst, size_t n) const {
    if (off > size || n > size - off) fail(MLF_ERR_IO, "read past the end of %s", path.c_str());
    uint8_t* d = static_cast<uint8_t*>(dst);
    while (n) {
      const ssize_t r = ::pread(fd, d, n, (off_t)off);
      if (r <= 0) fail(MLF_ERR_IO, "read error in %s", path.c_str());
      d += r;
      off += (uint64_t)r;
      n -= (size_t)r;
    }
  }
  uint64_t abs(uint64_t rel) const {
    if (rel == kUndef || rel > size || base > size - rel)
      fail(MLF_ERR_IO, "address %llu outside %s", (unsigned long long)rel, path.c_str());
    return base + rel;
  }
  void read(uint64_t rel, void* dst, size_t n) const { pread_abs(abs(rel), dst, n); }
  // n bytes at rel, checked against the file before anything is allocated
  std::vector<uint8_t> bytes(uint64_t rel, uint64_t n) const {
    const uint64_t a = abs(rel);
    if (n > size - a) fail(MLF_ERR_IO, "read past the end of %s", path.c_str());
    std::vector<uint8_t> b((size_t)n);
    pread_abs(a, b.data(), (size_t)n);
    return b;
  }
  std::vector<uint8_t> block(uint64_t rel, size_t n) const {
    if (n > kMaxMeta) fail(MLF_ERR_IO, "metadata block of %zu bytes", n);
    return bytes(rel, n);
  }
  // Longest block that fits before the end of the file, at most n bytes.
  std::vector<uint8_t> block_upto(uint64_t rel, size_t n) const {
    const uint64_t a = abs(rel);
    return block(rel, (size_t)std::min<uint64_t>(n, size - a));
  }
};

// ---------------------------------------------------------- object header
struct Msg {
  uint16_t type;
  uint8_t flags;
  std::vector<uint8_t> data;
};

enum : uint16_t {
  kDataspace = 0x01, kLinkInfo = 0x02, kDatatype = 0x03, kFillOld = 0x04, kFill = 0x05,
  kLink = 0x06, kLayout = 0x08, kFilters = 0x0B, kAttribute = 0x0C, kContinuation = 0x10,
  kSymbolTable = 0x11
};

static std::vector<Msg> object_header(const File& f, uint64_t addr) {
  std::vector<Msg> out;
  const std::vector<uint8_t> head = f.block_upto(addr, 16);
  if (head.size() < 16) fail(MLF_ERR_IO, "truncated object header");
  struct Pending { uint64_t addr, len; };
  std::vector<Pending> todo;
  std::set<uint64_t> seen_blocks;  // a continuation that loops back is malformed
  auto push = [&](uint64_t a, uint64_t l) {
    if (!seen_blocks.insert(a).second) fail(MLF_ERR_IO, "object header continuation loops");
    todo.push_back({a, l});
  };
  int blocks = 0;
  if (std::memcmp(head.data(), "OHDR", 4) == 0) {
    // version 2
    Cur c(head.data(), head.size(), "object header");
    c.skip(4);
    if (c.u8() != 2) fail(MLF_ERR_IO, "object header version");
    const uint8_t fl = c.u8();
    size_t pre = 6 + ((fl & 0x20) ? 16 : 0) + ((fl & 0x10) ? 4 : 0);
    const int wsz = 1 << (fl & 3);
    std::vector<uint8_t> h2 = f.block_upto(addr, pre + (size_t)wsz);
    Cur c2(h2.data(), h2.size(), "object header");
    c2.skip(pre);
    const uint64_t len0 = c2.un(wsz);
    const bool corder = fl & 0x04;
    auto parse = [&](uint64_t a, uint64_t len, bool first) {
      std::vector<uint8_t> b = f.block(a, (size_t)len);
      Cur m(b.data(), b.size(), "object header messages");
      if (!first) {
        if (len < 8 || std::memcmp(b.data(), "OCHK", 4) != 0) fail(MLF_ERR_IO, "bad continuation block");
        m.skip(4);
      }
      m.e -= first ? 0 : 4;  // trailing checksum of a continuation block
      const size_t mh = 4 + (corder ? 2 : 0);
      while (m.left() >= mh) {
        Msg msg;
        msg.type = m.u8();
        const uint16_t sz = m.u16();
        msg.flags = m.u8();
        if (corder) m.skip(2);
        const uint8_t* d = m.take(sz);
        msg.data.assign(d, d + sz);
        if (msg.type == kContinuation) {
          Cur cc(msg.data.data(), msg.data.size(), "continuation message");
          const uint64_t ca = cc.un(f.so), cl = cc.un(f.sl);
          push(ca, cl);
        } else if (msg.type != 0) {
          out.push_back(std::move(msg));
        }
      }
    };
    parse(addr + pre + (uint64_t)wsz, len0, true);
    while (!todo.empty()) {
      if (++blocks > 4096) fail(MLF_ERR_IO, "too many continuation blocks");
      const Pending p = todo.back();
      todo.pop_back();
      parse(p.addr, p.len, false);
    }
    return out;
  }
  // version 1: 12-byte prefix padded to 16, messages 8-byte aligned
  Cur c(head.data(), head.size(), "object header");
  if (c.u8() != 1) fail(MLF_ERR_IO, "unknown object header version at %llu", (unsigned long long)addr);
  c.skip(1);
  const uint16_t nmsg = c.u16();
  c.skip(4);
  const uint32_t hsize = c.u32();
  push(addr + 16, hsize);
  size_t seen = 0;
  while (!todo.empty() && seen < nmsg) {
    if (++blocks > 4096) fail(MLF_ERR_IO, "too many continuation blocks");
    const Pending p = todo.front();
    todo.erase(todo.begin());
    std::vector<uint8_t> b = f.block(p.addr, (size_t)p.len);
    Cur m(b.data(), b.size(), "object header messages");
    while (m.left() >= 8 && seen < nmsg) {
      Msg msg;
      msg.type = m.u16();
      const uint16_t sz = m.u16();
      msg.flags = m.u8();
      m.skip(3);
      const uint8_t* d = m.take(sz);
      msg.data.assign(d, d + sz);
      ++seen;
      if (msg.type == kContinuation) {
        Cur cc(msg.data.data(), msg.data.size(), "continuation message");
        const uint64_t ca = cc.un(f.so), cl = cc.un(f.sl);
        push(ca, cl);
      } else if (msg.type != 0) {
        out.push_back(std::move(msg));
      }
    }
  }
  return out;
}

static const Msg* find(const std::vector<Msg>& ms, uint16_t type) {
  for (const Msg& m : ms)
    if (m.type == type) {
      if (m.flags & 0x02) fail(MLF_ERR_UNSUPPORTED, "shared header message (type %d)", (int)type);
      return &m;
    }
  return nullptr;
}

// ----------------------------------------------------------- data types
struct Dtype {
  int cls = -1;        // 0 fixed-point, 1 float, 3 string, other: unsupported
  uint32_t size = 0;
  bool be = false;
  bool is_signed = false;
};

static Dtype parse_dtype(Cur& c) {
  Dtype t;
  const uint8_t cv = c.u8();
  t.cls = cv & 0x0F;
  const uint8_t b0 = c.u8();
  c.skip(2);
  t.size = c.u32();
  switch (t.cls) {
    case 0: {  // fixed-point
      t.be = b0 & 1;
      t.is_signed = (b0 >> 3) & 1;
      const uint16_t off = c.u16(), prec = c.u16();
      if (off != 0 || prec != 8 * t.size || (t.size != 1 && t.size != 2 && t.size != 4 && t.size != 8))
        fail(MLF_ERR_UNSUPPORTED, "integer type of %u bytes, %u bits at offset %u", t.size, prec, off);
      break;
    }
    case 1: {  // IEEE float
      if (b0 & 0x40) fail(MLF_ERR_UNSUPPORTED, "VAX float byte order");
      t.be = b0 & 1;
      const uint16_t off = c.u16(), prec = c.u16();
      const uint8_t eloc = c.u8(), esz = c.u8(), mloc = c.u8(), msz = c.u8();
      const uint32_t bias = c.u32();
      const bool f32 = t.size == 4 && eloc == 23 && esz == 8 && mloc == 0 && msz == 23 && bias == 127;
      const bool f64 = t.size == 8 && eloc == 52 && esz == 11 && mloc == 0 && msz == 52 && bias == 1023;
      if (off != 0 || prec != 8 * t.size || !(f32 || f64))
        fail(MLF_ERR_UNSUPPORTED, "float type of %u bytes is not IEEE single or double", t.size);
      break;
    }
    case 3:  // string
      break;
    case 6:
      fail(MLF_ERR_UNSUPPORTED, "compound type (MATLAB complex data)");
    case 7:
      fail(MLF_ERR_UNSUPPORTED, "reference type (MATLAB cell, struct array or object)");
    default:
      fail(MLF_ERR_UNSUPPORTED, "HDF5 datatype class %d", t.cls);
  }
  return t;
}

static std::vector<uint64_t> parse_dataspace(Cur& c, int sl) {
  const uint8_t ver = c.u8();
  const uint8_t rank = c.u8();
  const uint8_t flags = c.u8();
  std::vector<uint64_t> dims;
  if (ver == 1) {
    c.skip(5);
  } else if (ver == 2) {
    const uint8_t type = c.u8();
    if (type == 2) fail(MLF_ERR_UNSUPPORTED, "null dataspace");
  } else {
    fail(MLF_ERR_IO, "dataspace version %d", (int)ver);
  }
  if (rank > 32) fail(MLF_ERR_UNSUPPORTED, "rank %d", (int)rank);
  for (int i = 0; i < rank; ++i) dims.push_back(c.un(sl));
  if (flags & 1)
    for (int i = 0; i < rank; ++i) c.un(sl);
  return dims;
}

// ------------------------------------------------------------- attributes
struct Attr {
  std::string name;
  Dtype type;
  std::vector<uint64_t> dims;
  std::vector<uint8_t> value;
};

static Attr parse_attr(const File& f, const Msg& m) {
  Cur c(m.data.data(), m.data.size(), "attribute message");
  Attr a;
  const uint8_t ver = c.u8();
  if (ver < 1 || ver > 3) fail(MLF_ERR_IO, "attribute version %d", (int)ver);
  const uint8_t fl = c.u8();
  if (ver >= 2 && (fl & 3)) fail(MLF_ERR_UNSUPPORTED, "shared attribute type or space");
  const uint16_t nsz = c.u16(), tsz = c.u16(), ssz = c.u16();
  if (ver == 3) c.skip(1);
  auto pad = [&](size_t n) { return ver == 1 ? (n + 7) & ~size_t(7) : n; };
  const uint8_t* nm = c.take(pad(nsz));
  a.name.assign(reinterpret_cast<const char*>(nm), strnlen(reinterpret_cast<const char*>(nm), nsz));
  {
    const uint8_t* t = c.take(pad(tsz));
    Cur ct(t, tsz, "attribute datatype");
    try {
      a.type = parse_dtype(ct);
    } catch (const Error&) {
      a.type.cls = -1;  // an attribute this reader does not need
    }
  }
  {
    const uint8_t* s = c.take(pad(ssz));
    Cur cs(s, ssz, "attribute dataspace");
    a.dims = parse_dataspace(cs, f.sl);
  }
  a.value.assign(c.p, c.e);
  return a;
}

// --------------------------------------------------------------- groups
static std::string heap_string(const std::vector<uint8_t>& heap, uint64_t off) {
  if (off >= heap.size()) fail(MLF_ERR_IO, "name offset outside the local heap");
  const char* s = reinterpret_cast<const char*>(heap.data() + off);
  return std::string(s, strnlen(s, heap.size() - off));
}

static void symbol_btree(const File& f, uint64_t addr, const std::vector<uint8_t>& heap, int want_level,
                         std::set<uint64_t>* seen, std::vector<std::pair<std::string, uint64_t>>* out) {
  if (!seen->insert(addr).second || seen->size() > (1u << 20)) fail(MLF_ERR_IO, "group B-tree loops");
  const size_t hdr = 8 + 2 * (size_t)f.so;
  std::vector<uint8_t> h = f.block(addr, hdr);
  Cur c(h.data(), h.size(), "group B-tree node");
  if (std::memcmp(c.take(4), "TREE", 4) != 0) fail(MLF_ERR_IO, "bad group B-tree signature");
  if (c.u8() != 0) fail(MLF_ERR_IO, "group B-tree of the wrong node type");
  const uint8_t level = c.u8();
  if (want_level >= 0 && level != want_level) fail(MLF_ERR_IO, "group B-tree level");
  const uint16_t n = c.u16();
  const size_t body = (size_t)n * (f.sl + f.so) + f.sl;
  std::vector<uint8_t> b = f.block(addr + hdr, body);
  Cur k(b.data(), b.size(), "group B-tree node");
  for (int i = 0; i < n; ++i) {
    k.un(f.sl);
    const uint64_t child = k.un(f.so);
    if (level > 0) {
      symbol_btree(f, child, heap, level - 1, seen, out);
      continue;
    }
    std::vector<uint8_t> sh = f.block(child, 8);
    Cur s(sh.data(), sh.size(), "symbol table node");
    if (std::memcmp(s.take(4), "SNOD", 4) != 0) fail(MLF_ERR_IO, "bad symbol table node");
    s.skip(2);
    const uint16_t ns = s.u16();
    const size_t esz = 2 * (size_t)f.so + 24;
    std::vector<uint8_t> eb = f.block(child + 8, (size_t)ns * esz);
    Cur e(eb.data(), eb.size(), "symbol table entries");
    for (int j = 0; j < ns; ++j) {
      const uint64_t name_off = e.un(f.so);
      const uint64_t oh = e.un(f.so);
      e.skip(24);
      out->emplace_back(heap_string(heap, name_off), oh);
    }
  }
}

// Children (name, object header address) of the group at `addr`.
static std::vector<std::pair<std::string, uint64_t>> group_members(const File& f, uint64_t addr) {
  const std::vector<Msg> ms = object_header(f, addr);
  std::vector<std::pair<std::string, uint64_t>> out;
  if (const Msg* st = find(ms, kSymbolTable)) {
    Cur c(st->data.data(), st->data.size(), "symbol table message");
    const uint64_t bt = c.un(f.so), hp = c.un(f.so);
    std::vector<uint8_t> hh = f.block(hp, 8 + 2 * (size_t)f.sl + f.so);
    Cur h(hh.data(), hh.size(), "local heap");
    if (std::memcmp(h.take(4), "HEAP", 4) != 0) fail(MLF_ERR_IO, "bad local heap");
    h.skip(4);
    const uint64_t dsize = h.un(f.sl);
    h.un(f.sl);
    const uint64_t daddr = h.un(f.so);
    const std::vector<uint8_t> heap = f.block(daddr, (size_t)dsize);
    std::set<uint64_t> seen;
    symbol_btree(f, bt, heap, -1, &seen, &out);
    return out;
  }
  if (const Msg* li = find(ms, kLinkInfo)) {
    Cur c(li->data.data(), li->data.size(), "link info message");
    c.skip(1);
    const uint8_t fl = c.u8();
    if (fl & 1) c.skip(8);
    if (c.un(f.so) != kUndef) fail(MLF_ERR_UNSUPPORTED, "dense link storage (fractal heap)");
  }
  for (const Msg& m : ms) {
    if (m.type != kLink) continue;
    Cur c(m.data.data(), m.data.size(), "link message");
    if (c.u8() != 1) fail(MLF_ERR_IO, "link message version");
    const uint8_t fl = c.u8();
    const uint8_t ltype = (fl & 0x08) ? c.u8() : 0;
    if (fl & 0x04) c.skip(8);
    if (fl & 0x10) c.skip(1);
    const uint64_t nlen = c.un(1 << (fl & 3));
    const uint8_t* nm = c.take((size_t)nlen);
    if (ltype != 0) continue;  // soft and external links are not followed
    out.emplace_back(std::string(reinterpret_cast<const char*>(nm), (size_t)nlen), c.un(f.so));
  }
  return out;
}

// ---------------------------------------------------------------- dataset
struct Filter {
  uint16_t id;
  std::vector<uint32_t> cd;
};

struct Chunk {
  uint64_t addr, size;     // relative address, stored bytes
  uint32_t mask;           // filters skipped for this chunk
  std::vector<uint64_t> off;  // element offsets, HDF5 order
};

struct Dataset {
  std::string name;
  std::string mclass;  // MATLAB_class ("" when the attribute is absent)
  bool empty = false;  // MATLAB_empty: the data holds the MATLAB dims
  bool sparse = false;
  bool is_group = false;
  Dtype type;
  std::vector<uint64_t> dims;  // HDF5 order
  int layout = -1;             // 0 compact, 1 contiguous, 2 chunked
  std::vector<uint8_t> compact;
  uint64_t addr = kUndef, csize = 0;
  std::vector<uint64_t> cdims;  // chunk dims, HDF5 order (rank entries)
  int index = 0;                // chunked: 0 v1 B-tree, 1 single, 2 implicit, 3 fixed array
  uint64_t single_size = 0;
  uint32_t single_mask = 0;
  bool single_filtered = false;
  std::vector<Filter> filters;
  std::vector<uint8_t> fill;  // one element, or empty for zeros
  std::string type_error;     // why the datatype cannot be read

  uint64_t numel() const {
    uint64_t n = 1;
    for (uint64_t d : dims) {
      if (d && n > (~0ull >> 1) / d) fail(MLF_ERR_IO, "dataset size overflows");
      n *= d;
    }
    return n;
  }
};

static void parse_layout(const File& f, const Msg& m, Dataset* d) {
  Cur c(m.data.data(), m.data.size(), "layout message");
  const uint8_t ver = c.u8();
  if (ver < 1 || ver > 4) fail(MLF_ERR_UNSUPPORTED, "layout message version %d", (int)ver);
  if (ver < 3) {
    // versions 1 and 2: rank, class, 5 reserved bytes, address (not for
    // compact), rank 32-bit dims (chunked: chunk dims + element size)
    const uint8_t r = c.u8();
    d->layout = c.u8();
    c.skip(5);
    if (d->layout == 1 || d->layout == 2) d->addr = c.un(f.so);
    std::vector<uint64_t> dims;
    for (int i = 0; i < r; ++i) dims.push_back(c.u32());
    if (d->layout == 0) {
      const uint32_t n = c.u32();
      const uint8_t* p = c.take(n);
      d->compact.assign(p, p + n);
    } else if (d->layout == 1) {
      d->csize = kUndef;  // implied by the dataspace
    } else if (d->layout == 2) {
      if (r < 1) fail(MLF_ERR_IO, "chunk rank");
      d->cdims = dims;
      d->cdims.pop_back();
      d->index = 0;
    } else {
      fail(MLF_ERR_UNSUPPORTED, "layout class %d", d->layout);
    }
    return;
  }
  d->layout = c.u8();
  if (d->layout == 0) {
    const uint16_t n = c.u16();
    const uint8_t* p = c.take(n);
    d->compact.assign(p, p + n);
  } else if (d->layout == 1) {
    d->addr = c.un(f.so);
    d->csize = c.un(f.sl);
  } else if (d->layout == 2) {
    if (ver == 3) {
      const uint8_t r = c.u8();
      d->addr = c.un(f.so);
      if (r < 1) fail(MLF_ERR_IO, "chunk rank");
      for (int i = 0; i < r; ++i) d->cdims.push_back(c.u32());
      d->index = 0;
    } else {
      const uint8_t fl = c.u8();
      const uint8_t r = c.u8();
      const uint8_t enc = c.u8();
      if (r < 1 || enc < 1 || enc > 8) fail(MLF_ERR_IO, "chunk layout");
      for (int i = 0; i < r; ++i) d->cdims.push_back(c.un(enc));
      const uint8_t it = c.u8();
      if (it == 1) {
        d->index = 1;
        if (fl & 2) {
          d->single_filtered = true;
          d->single_size = c.un(f.sl);
          d->single_mask = c.u32();
        }
      } else if (it == 2) {
        d->index = 2;
      } else if (it == 3) {
        d->index = 3;
        c.skip(1);
      } else {
        fail(MLF_ERR_UNSUPPORTED, "chunk index type %d (extensible array / v2 B-tree)", (int)it);
      }
      d->addr = c.un(f.so);
    }
    d->cdims.pop_back();  // the last chunk dimension is the element size
  } else {
    fail(MLF_ERR_UNSUPPORTED, "layout class %d", d->layout);
  }
}

static void parse_filters(const Msg& m, Dataset* d) {
  Cur c(m.data.data(), m.data.size(), "filter pipeline message");
  const uint8_t ver = c.u8();
  const uint8_t n = c.u8();
  if (ver == 1) c.skip(6);
  else if (ver != 2) fail(MLF_ERR_IO, "filter pipeline version %d", (int)ver);
  for (int i = 0; i < n; ++i) {
    Filter flt;
    flt.id = c.u16();
    uint16_t nlen = 0;
    if (ver == 1 || flt.id >= 256) nlen = c.u16();
    c.u16();  // flags
    const uint16_t nv = c.u16();
    if (nlen) c.skip(ver == 1 ? ((size_t)nlen + 7) & ~size_t(7) : nlen);
    for (int k = 0; k < nv; ++k) flt.cd.push_back(c.u32());
    if (ver == 1 && (nv & 1)) c.skip(4);
    if (flt.id != 1 && flt.id != 2 && flt.id != 3)
      fail(MLF_ERR_UNSUPPORTED, "HDF5 filter %d (only deflate, shuffle and fletcher32 are decoded)",
           (int)flt.id);
    d->filters.push_back(std::move(flt));
  }
}

static Dataset describe(const File& f, const std::string& name, uint64_t addr) {
  Dataset d;
  d.name = name;
  const std::vector<Msg> ms = object_header(f, addr);
  for (const Msg& m : ms) {
    if (m.type != kAttribute) continue;
    const Attr a = parse_attr(f, m);
    if (a.name == "MATLAB_class" && a.type.cls == 3) {
      d.mclass.assign(reinterpret_cast<const char*>(a.value.data()),
                      strnlen(reinterpret_cast<const char*>(a.value.data()),
                              std::min<size_t>(a.value.size(), a.type.size)));
    } else if (a.name == "MATLAB_empty" && a.type.cls == 0 && !a.value.empty()) {
      for (uint8_t b : a.value) d.empty |= b != 0;
    } else if (a.name == "MATLAB_sparse") {
      d.sparse = true;
    }
  }
  if (find(ms, kSymbolTable) || find(ms, kLinkInfo) || (!find(ms, kLayout) && find(ms, kLink))) {
    d.is_group = true;
    return d;
  }
  const Msg* sp = find(ms, kDataspace);
  const Msg* ty = find(ms, kDatatype);
  const Msg* ly = find(ms, kLayout);
  if (!sp || !ty || !ly) fail(MLF_ERR_IO, "variable %s has no dataspace, datatype or layout", name.c_str());
  {
    Cur c(sp->data.data(), sp->data.size(), "dataspace message");
    d.dims = parse_dataspace(c, f.sl);
  }
  try {
    Cur c(ty->data.data(), ty->data.size(), "datatype message");
    d.type = parse_dtype(c);
  } catch (const Error& e) {
    if (e.code != MLF_ERR_UNSUPPORTED) throw;
    d.type.cls = -1;  // listed by mlf_mat_info, refused by the readers
    d.type_error = e.what();
  }
  parse_layout(f, *ly, &d);
  if (const Msg* fp = find(ms, kFilters)) parse_filters(*fp, &d);
  if (const Msg* fv = find(ms, kFill)) {
    Cur c(fv->data.data(), fv->data.size(), "fill value message");
    const uint8_t ver = c.u8();
    bool defined = false;
    if (ver == 1 || ver == 2) {
      c.skip(2);
      defined = c.u8() != 0;
      if (ver == 1) defined = true;  // v1 always carries the size field
    } else if (ver == 3) {
      defined = (c.u8() & 0x20) != 0;
    } else {
      fail(MLF_ERR_IO, "fill value version %d", (int)ver);
    }
    if (defined && c.left() >= 4) {
      const uint32_t n = c.u32();
      if (n && n == d.type.size) {
        const uint8_t* p = c.take(n);
        d.fill.assign(p, p + n);
      }
    }
  }
  if (d.layout == 2 && d.cdims.size() != d.dims.size())
    fail(MLF_ERR_IO, "variable %s: chunk rank %zu for data rank %zu", name.c_str(), d.cdims.size(),
         d.dims.size());
  return d;
}

// ------------------------------------------------------------ conversion
template <typename Out>
static void convert_run(const uint8_t* src, const Dtype& t, size_t n, Out* dst) {
  const uint32_t es = t.size;
  auto load = [&](const uint8_t* p) -> uint64_t {
    uint64_t v = 0;
    if (t.be)
      for (uint32_t i = 0; i < es; ++i) v = (v << 8) | p[i];
    else
      for (uint32_t i = 0; i < es; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
  };
  if (t.cls == 1) {
    if (es == 4) {
      for (size_t i = 0; i < n; ++i) {
        const uint32_t b = (uint32_t)load(src + i * 4);
        float x;
        std::memcpy(&x, &b, 4);
        dst[i] = (Out)x;
      }
    } else {
      for (size_t i = 0; i < n; ++i) {
        const uint64_t b = load(src + i * 8);
        double x;
        std::memcpy(&x, &b, 8);
        dst[i] = (Out)x;
      }
    }
    return;
  }
  for (size_t i = 0; i < n; ++i) {
    const uint64_t v = load(src + i * es);
    if (t.is_signed) {
      const int sh = 64 - 8 * (int)es;
      dst[i] = (Out)(int64_t)((int64_t)(v << sh) >> sh);
    } else {
      dst[i] = (Out)v;
    }
  }
}

// -------------------------------------------------------------- filters
// Inflate a zlib stream of at most `limit` bytes of output (deflate expands
// by at most 1032:1, which bounds the first allocation).
static std::vector<uint8_t> inflate_all(const std::vector<uint8_t>& in, size_t limit) {
  std::vector<uint8_t> out(std::min<size_t>(limit, in.size() * 1032 + 1024));
  z_stream zs;
  std::memset(&zs, 0, sizeof(zs));
  if (inflateInit(&zs) != Z_OK) fail(MLF_ERR_INTERNAL, "inflateInit failed");
  zs.next_in = const_cast<Bytef*>(in.data());
  zs.avail_in = (uInt)in.size();
  size_t done = 0;
  int rc = Z_OK;
  while (rc != Z_STREAM_END) {
    if (done == out.size()) {
      if (out.size() >= limit) {
        inflateEnd(&zs);
        fail(MLF_ERR_IO, "deflate stream longer than its chunk");
      }
      out.resize(std::min<size_t>(limit, out.size() * 2));
    }
    const size_t room = std::min<size_t>(out.size() - done, 1u << 30);
    zs.next_out = out.data() + done;
    zs.avail_out = (uInt)room;
    rc = inflate(&zs, Z_NO_FLUSH);
    done += room - zs.avail_out;
    if (rc == Z_BUF_ERROR && zs.avail_in == 0) break;
    if (rc != Z_OK && rc != Z_STREAM_END && rc != Z_BUF_ERROR) {
      inflateEnd(&zs);
      fail(MLF_ERR_IO, "corrupt deflate stream (zlib %d)", rc);
    }
  }
  inflateEnd(&zs);
  if (rc != Z_STREAM_END) fail(MLF_ERR_IO, "truncated deflate stream");
  out.resize(done);
  return out;
}

static std::vector<uint8_t> unshuffle(const std::vector<uint8_t>& in, size_t es) {
  if (es <= 1) return in;
  std::vector<uint8_t> out(in.size());
  const size_t n = in.size() / es;
  for (size_t b = 0; b < es; ++b)
    for (size_t i = 0; i < n; ++i) out[i * es + b] = in[b * n + i];
  std::memcpy(out.data() + n * es, in.data() + n * es, in.size() - n * es);
  return out;
}

static std::vector<uint8_t> decode_chunk(const File& f, const Dataset& d, const Chunk& ch, size_t raw) {
  std::vector<uint8_t> buf = f.bytes(ch.addr, ch.size);
  for (size_t k = d.filters.size(); k-- > 0;) {
    if (ch.mask & (1u << k)) continue;
    const Filter& flt = d.filters[k];
    if (flt.id == 1) {
      buf = inflate_all(buf, raw + 64);
    } else if (flt.id == 2) {
      buf = unshuffle(buf, flt.cd.empty() ? d.type.size : flt.cd[0]);
    } else if (flt.id == 3) {
      if (buf.size() < 4) fail(MLF_ERR_IO, "fletcher32 chunk too short");
      buf.resize(buf.size() - 4);
    }
  }
  if (buf.size() < raw) fail(MLF_ERR_IO, "chunk of %s decodes to %zu bytes, expected %zu",
                             d.name.c_str(), buf.size(), raw);
  return buf;
}

// ---------------------------------------------------------- chunk lists
static void chunk_btree(const File& f, const Dataset& d, uint64_t addr, int want_level, std::set<uint64_t>* seen,
                        std::vector<Chunk>* out) {
  if (!seen->insert(addr).second || seen->size() > (1u << 22)) fail(MLF_ERR_IO, "chunk B-tree loops");
  const size_t hdr = 8 + 2 * (size_t)f.so;
  std::vector<uint8_t> h = f.block(addr, hdr);
  Cur c(h.data(), h.size(), "chunk B-tree node");
  if (std::memcmp(c.take(4), "TREE", 4) != 0) fail(MLF_ERR_IO, "bad chunk B-tree signature");
  if (c.u8() != 1) fail(MLF_ERR_IO, "chunk B-tree of the wrong node type");
  const uint8_t level = c.u8();
  if (want_level >= 0 && level != want_level) fail(MLF_ERR_IO, "chunk B-tree level");
  const uint16_t n = c.u16();
  const size_t r = d.dims.size();
  const size_t ksz = 8 + 8 * (r + 1);
  std::vector<uint8_t> b = f.block(addr + hdr, (size_t)n * (ksz + f.so) + ksz);
  Cur k(b.data(), b.size(), "chunk B-tree node");
  for (int i = 0; i < n; ++i) {
    Chunk ch;
    ch.size = k.u32();
    ch.mask = k.u32();
    for (size_t j = 0; j < r; ++j) ch.off.push_back(k.u64());
    k.u64();
    ch.addr = k.un(f.so);
    if (level > 0) {
      chunk_btree(f, d, ch.addr, level - 1, seen, out);
    } else {
      out->push_back(std::move(ch));
    }
  }
}

// Bytes of one unfiltered chunk (HDF5 caps chunks at 4 GiB).
static uint64_t chunk_bytes(const Dataset& d) {
  uint64_t raw = d.type.size;
  for (uint64_t c : d.cdims) {
    if (c == 0) fail(MLF_ERR_IO, "zero chunk dimension");
    if (raw > (1ull << 32) / c) fail(MLF_ERR_IO, "chunk larger than 4 GiB");
    raw *= c;
  }
  return raw;
}

// Chunks in the grid over the extent.
static uint64_t grid_count(const Dataset& d) {
  uint64_t total = 1;
  for (size_t j = 0; j < d.dims.size(); ++j) {
    const uint64_t n = d.dims[j] / d.cdims[j] + (d.dims[j] % d.cdims[j] != 0);
    if (n && total > (1ull << 40) / n) fail(MLF_ERR_IO, "chunk grid too large");
    total *= n;
  }
  return total;
}

// Offsets of chunk i of the row-major chunk grid, the order the implicit and
// fixed-array indexes store chunks in.
static std::vector<uint64_t> grid_offset(const Dataset& d, uint64_t i) {
  const size_t r = d.dims.size();
  std::vector<uint64_t> off(r);
  for (size_t j = r; j-- > 0;) {
    const uint64_t n = d.dims[j] / d.cdims[j] + (d.dims[j] % d.cdims[j] != 0);
    off[j] = (i % n) * d.cdims[j];
    i /= n;
  }
  return off;
}

static std::vector<Chunk> chunk_list(const File& f, const Dataset& d) {
  std::vector<Chunk> out;
  if (d.addr == kUndef) return out;  // nothing written: all fill value
  const uint64_t raw = chunk_bytes(d);
  if (d.index == 0) {
    std::set<uint64_t> seen;
    chunk_btree(f, d, d.addr, -1, &seen, &out);
  } else if (d.index == 1) {
    Chunk ch;
    ch.addr = d.addr;
    ch.size = d.single_filtered ? d.single_size : raw;
    ch.mask = d.single_filtered ? d.single_mask : 0;
    ch.off.assign(d.dims.size(), 0);
    out.push_back(ch);
  } else if (d.index == 2) {
    // unfiltered chunks back to back from addr
    const uint64_t total = grid_count(d);
    const uint64_t a = f.abs(d.addr);
    if (total > (f.size - a) / raw) fail(MLF_ERR_IO, "implicit chunk index runs past the end of the file");
    for (uint64_t i = 0; i < total; ++i) out.push_back({d.addr + i * raw, raw, 0, grid_offset(d, i)});
  } else {
    std::vector<uint8_t> hh = f.block(d.addr, 12 + (size_t)f.sl + f.so);
    Cur h(hh.data(), hh.size(), "fixed array header");
    if (std::memcmp(h.take(4), "FAHD", 4) != 0) fail(MLF_ERR_IO, "bad fixed array header");
    h.skip(1);
    const uint8_t client = h.u8();
    const uint8_t esz = h.u8();
    const uint8_t page_bits = h.u8();
    const uint64_t nent = h.un(f.sl);
    const uint64_t db = h.un(f.so);
    if (nent > (1ull << page_bits)) fail(MLF_ERR_UNSUPPORTED, "paged fixed array chunk index");
    const uint64_t total = grid_count(d);
    if (total != nent || esz < f.so) fail(MLF_ERR_IO, "fixed array holds %llu chunks, grid has %llu",
                                          (unsigned long long)nent, (unsigned long long)total);
    std::vector<uint8_t> bb = f.block(db, 6 + (size_t)f.so + (size_t)nent * esz);
    Cur b(bb.data(), bb.size(), "fixed array data block");
    if (std::memcmp(b.take(4), "FADB", 4) != 0) fail(MLF_ERR_IO, "bad fixed array data block");
    b.skip(2 + f.so);
    for (uint64_t i = 0; i < nent; ++i) {
      Chunk ch;
      ch.addr = b.un(f.so);
      if (client == 1) {
        const int szb = (int)esz - f.so - 4;
        if (szb < 1 || szb > 8) fail(MLF_ERR_IO, "fixed array entry size");
        ch.size = b.un(szb);
        ch.mask = b.u32();
      } else {
        ch.size = raw;
        ch.mask = 0;
      }
      if (client != 1) b.skip((size_t)esz - f.so);
      ch.off = grid_offset(d, i);
      if (ch.addr != kUndef) out.push_back(std::move(ch));
    }
  }
  return out;
}

// -------------------------------------------------------------- reading
template <typename Out>
static void fill_all(const Dataset& d, Out* out, uint64_t n) {
  Out v = 0;
  if (!d.fill.empty()) convert_run(d.fill.data(), d.type, 1, &v);
  for (uint64_t i = 0; i < n; ++i) out[i] = v;
}

static void run_parallel(size_t n, const std::function<void(size_t)>& fn) {
  const size_t hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min<size_t>(hw, 32), n);
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::string err;
  mlf_status code = MLF_OK;
  std::mutex mu;
  std::vector<std::thread> ts;
  for (size_t t = 0; t < nt; ++t)
    ts.emplace_back([&] {
      for (;;) {
        const size_t i = next.fetch_add(1);
        if (i >= n) return;
        try {
          fn(i);
        } catch (const Error& e) {
          std::lock_guard<std::mutex> g(mu);
          if (code == MLF_OK) { code = e.code; err = e.what(); }
          next.store(n);
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> g(mu);
          if (code == MLF_OK) { code = MLF_ERR_OUT_OF_MEMORY; err = e.what(); }
          next.store(n);
        }
      }
    });
  for (auto& t : ts) t.join();
  if (code != MLF_OK) throw Error(code, err);
}

template <typename Out>
static void read_data(const File& f, const Dataset& d, Out* out) {
  const uint64_t n = d.numel();
  const uint32_t es = d.type.size;
  if (n == 0) return;
  if (d.layout == 0) {
    if (n > d.compact.size() / es) fail(MLF_ERR_IO, "compact data of %s too short", d.name.c_str());
    convert_run(d.compact.data(), d.type, (size_t)n, out);
    return;
  }
  if (d.layout == 1) {
    if (d.addr == kUndef) return fill_all(d, out, n);
    if (n > f.size / es || (d.csize != kUndef && d.csize / es < n))
      fail(MLF_ERR_IO, "contiguous data of %s too short", d.name.c_str());
    const uint64_t slab = (16ull << 20) / es;
    const size_t nslab = (size_t)((n + slab - 1) / slab);
    run_parallel(nslab, [&](size_t s) {
      const uint64_t a = s * slab, m = std::min<uint64_t>(slab, n - a);
      std::vector<uint8_t> buf((size_t)(m * es));
      f.read(d.addr + a * es, buf.data(), buf.size());
      convert_run(buf.data(), d.type, (size_t)m, out + a);
    });
    return;
  }
  const size_t r = d.dims.size();
  const uint64_t cn = chunk_bytes(d) / es;
  const std::vector<Chunk> chunks = chunk_list(f, d);
  // Chunks cover the extent exactly when every grid cell is present; otherwise
  // the holes take the fill value.
  if (chunks.size() < grid_count(d)) fill_all(d, out, n);
  std::vector<uint64_t> stride(r, 1);
  for (size_t j = r - 1; j-- > 0;) stride[j] = stride[j + 1] * d.dims[j + 1];
  run_parallel(chunks.size(), [&](size_t i) {
    const Chunk& ch = chunks[i];
    if (ch.off.size() != r) fail(MLF_ERR_IO, "chunk rank");
    for (size_t j = 0; j < r; ++j)
      if (ch.off[j] >= d.dims[j] || ch.off[j] % d.cdims[j]) fail(MLF_ERR_IO, "chunk outside %s", d.name.c_str());
    const std::vector<uint8_t> buf = decode_chunk(f, d, ch, (size_t)(cn * es));
    // rows along the fastest dimension, clipped to the extent
    std::vector<uint64_t> ext(r);
    for (size_t j = 0; j < r; ++j) ext[j] = std::min(d.cdims[j], d.dims[j] - ch.off[j]);
    const uint64_t run = ext[r - 1];
    uint64_t rows = 1;
    for (size_t j = 0; j + 1 < r; ++j) rows *= ext[j];
    std::vector<uint64_t> idx(r, 0);
    for (uint64_t row = 0; row < rows; ++row) {
      uint64_t src = 0, dst = 0;
      for (size_t j = 0; j < r; ++j) {
        uint64_t cs = 1;
        for (size_t k = j + 1; k < r; ++k) cs *= d.cdims[k];
        src += idx[j] * cs;
        dst += (ch.off[j] + idx[j]) * stride[j];
      }
      convert_run(buf.data() + src * es, d.type, (size_t)run, out + dst);
      for (size_t j = r - 1; j-- > 0;) {
        if (++idx[j] < ext[j]) break;
        idx[j] = 0;
      }
    }
  });
}

}  // namespace mat73
}  // namespace mlf

using namespace mlf::mat73;

struct mlf_mat_s {
  File file;
  std::vector<std::pair<std::string, uint64_t>> vars;  // root group members
};

namespace {

mlf_status guard(const std::function<void()>& fn) {
  try {
    fn();
    return MLF_OK;
  } catch (const Error& e) {
    return mlf::set_error(e.code, "%s", e.what());
  } catch (const std::bad_alloc&) {
    return mlf::set_error(MLF_ERR_OUT_OF_MEMORY, "host allocation failed reading a MAT-file");
  } catch (const std::exception& e) {
    return mlf::set_error(MLF_ERR_INTERNAL, "%s", e.what());
  }
}

Dataset lookup(const mlf_mat_s* m, const char* name) {
  if (!name) fail(MLF_ERR_INVALID_ARGUMENT, "name is NULL");
  // '/'-separated paths reach into groups (MATLAB structs)
  std::string path(name);
  const std::vector<std::pair<std::string, uint64_t>>* level = &m->vars;
  std::vector<std::pair<std::string, uint64_t>> sub;
  size_t pos = 0;
  for (int depth = 0;; ++depth) {
    if (depth > kMaxDepth) fail(MLF_ERR_INVALID_ARGUMENT, "path too deep");
    const size_t slash = path.find('/', pos);
    const std::string part = path.substr(pos, slash == std::string::npos ? std::string::npos : slash - pos);
    uint64_t addr = kUndef;
    for (const auto& v : *level)
      if (v.first == part) addr = v.second;
    if (addr == kUndef)
      fail(MLF_ERR_INVALID_ARGUMENT, "variable '%s' not found in %s", name, m->file.path.c_str());
    if (slash == std::string::npos) return describe(m->file, path, addr);
    sub = group_members(m->file, addr);
    level = &sub;
    pos = slash + 1;
  }
}

bool numeric_class(const std::string& c) {
  static const char* const ok[] = {"double", "single", "int8", "uint8", "int16", "uint16", "int32",
                                   "uint32", "int64", "uint64", "logical", "char"};
  for (const char* k : ok)
    if (c == k) return true;
  return false;
}

void check_readable(const Dataset& d) {
  if (d.is_group)
    fail(MLF_ERR_UNSUPPORTED, "variable '%s' is a MATLAB %s (a group), not a numeric array", d.name.c_str(),
         d.mclass.empty() ? "group" : d.mclass.c_str());
  if (d.sparse) fail(MLF_ERR_UNSUPPORTED, "variable '%s' is sparse", d.name.c_str());
  if (!d.mclass.empty() && !numeric_class(d.mclass))
    fail(MLF_ERR_UNSUPPORTED, "variable '%s' has MATLAB class %s", d.name.c_str(), d.mclass.c_str());
  if (!d.type_error.empty()) fail(MLF_ERR_UNSUPPORTED, "variable '%s': %s", d.name.c_str(), d.type_error.c_str());
  if (d.type.cls != 0 && d.type.cls != 1)
    fail(MLF_ERR_UNSUPPORTED, "variable '%s' is not numeric", d.name.c_str());
}

void info_of(const File& f, const Dataset& d, mlf_mat_var* info) {
  std::memset(info, 0, sizeof(*info));
  std::snprintf(info->matlab_class, sizeof(info->matlab_class), "%s",
                d.is_group ? (d.mclass.empty() ? "group" : d.mclass.c_str()) : d.mclass.c_str());
  if (d.is_group) return;
  info->numeric = (d.type.cls == 0 || d.type.cls == 1) && !d.sparse &&
                  (d.mclass.empty() || numeric_class(d.mclass));
  info->stored_bytes = (int32_t)d.type.size;
  if (d.empty) {
    // the data of an empty variable lists its MATLAB dims
    info->numel = 0;
    const uint64_t k = d.numel();
    if (d.type.cls == 0 && k >= 2 && k <= MLF_MAT_MAX_DIMS) {
      std::vector<double> v((size_t)k);
      read_data(f, d, v.data());
      info->ndim = (int32_t)k;
      for (uint64_t j = 0; j < k; ++j) info->dims[j] = (int64_t)v[(size_t)j];
    } else {
      info->ndim = 2;
    }
    return;
  }
  const size_t r = d.dims.size();
  if (r > MLF_MAT_MAX_DIMS) fail(MLF_ERR_UNSUPPORTED, "rank %zu", r);
  info->ndim = r == 0 ? 2 : (int32_t)r;
  if (r == 0) {
    info->dims[0] = info->dims[1] = 1;
  } else {
    for (size_t j = 0; j < r; ++j) info->dims[j] = (int64_t)d.dims[r - 1 - j];
    if (r == 1) {
      info->ndim = 2;
      info->dims[1] = 1;
    }
  }
  info->numel = (int64_t)d.numel();
}

}  // namespace

extern "C" {

mlf_status mlf_mat_open(const char* path, mlf_mat* out) {
  mlf::clear_error();
  if (!path || !out) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  std::unique_ptr<mlf_mat_s> m(new mlf_mat_s);
  const mlf_status st = guard([&] {
    File& f = m->file;
    f.path = path;
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) fail(MLF_ERR_IO, "cannot open %s", path);
    struct stat sb;
    if (fstat(f.fd, &sb) != 0) fail(MLF_ERR_IO, "cannot stat %s", path);
    f.size = (uint64_t)sb.st_size;
    // the superblock sits at 0 or at a power of two >= 512 (after a user block)
    static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    uint64_t at = kUndef;
    for (uint64_t o = 0; o + 8 <= f.size; o = o ? o * 2 : 512) {
      uint8_t s[8];
      f.pread_abs(o, s, 8);
      if (std::memcmp(s, sig, 8) == 0) {
        at = o;
        break;
      }
    }
    if (at == kUndef) fail(MLF_ERR_IO, "%s is not an HDF5 (MATLAB -v7.3) file", path);
    uint8_t sbuf[128] = {0};
    const size_t sn = (size_t)std::min<uint64_t>(sizeof(sbuf), f.size - at);
    f.pread_abs(at, sbuf, sn);
    Cur c(sbuf, sn, "superblock");
    c.skip(8);
    const uint8_t ver = c.u8();
    f.base = at;
    if (ver == 0 || ver == 1) {
      c.skip(4);
      f.so = c.u8();
      f.sl = c.u8();
      if ((f.so != 4 && f.so != 8) || (f.sl != 4 && f.sl != 8)) fail(MLF_ERR_UNSUPPORTED, "offset size");
      c.skip(1 + 4 + 4);
      if (ver == 1) c.skip(4);
      c.un(f.so);  // base address field: the library keeps it equal to the superblock's offset
      c.un(f.so);
      c.un(f.so);
      c.un(f.so);
      c.un(f.so);  // root symbol table entry: link name offset
      f.root = c.un(f.so);
    } else if (ver == 2 || ver == 3) {
      f.so = c.u8();
      f.sl = c.u8();
      if ((f.so != 4 && f.so != 8) || (f.sl != 4 && f.sl != 8)) fail(MLF_ERR_UNSUPPORTED, "offset size");
      c.skip(1);
      c.un(f.so);
      c.un(f.so);
      c.un(f.so);
      f.root = c.un(f.so);
    } else {
      fail(MLF_ERR_UNSUPPORTED, "HDF5 superblock version %d", (int)ver);
    }
    if (f.so == 4 && f.root == 0xFFFFFFFFull) f.root = kUndef;
    m->vars = group_members(f, f.root);
  });
  if (st != MLF_OK) return st;
  *out = m.release();
  return MLF_OK;
}

mlf_status mlf_mat_close(mlf_mat m) {
  mlf::clear_error();
  delete m;
  return MLF_OK;
}

mlf_status mlf_mat_count(mlf_mat m, int32_t* n) {
  mlf::clear_error();
  if (!m || !n) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *n = (int32_t)m->vars.size();
  return MLF_OK;
}

mlf_status mlf_mat_name(mlf_mat m, int32_t i, char* name, size_t cap) {
  mlf::clear_error();
  if (!m || !name || cap == 0) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (i < 0 || (size_t)i >= m->vars.size())
    return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "variable index %d out of range", i);
  const std::string& s = m->vars[(size_t)i].first;
  if (s.size() + 1 > cap) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "name buffer too small");
  std::memcpy(name, s.c_str(), s.size() + 1);
  return MLF_OK;
}

mlf_status mlf_mat_info(mlf_mat m, const char* name, mlf_mat_var* info) {
  mlf::clear_error();
  if (!m || !info) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  return guard([&] { info_of(m->file, lookup(m, name), info); });
}

mlf_status mlf_mat_read_f32(mlf_mat m, const char* name, float* out, int64_t numel) {
  mlf::clear_error();
  if (!m || (!out && numel)) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  return guard([&] {
    const Dataset d = lookup(m, name);
    check_readable(d);
    const int64_t n = d.empty ? 0 : (int64_t)d.numel();
    if (n != numel)
      fail(MLF_ERR_INVALID_ARGUMENT, "variable '%s' has %lld elements, buffer holds %lld", name,
           (long long)n, (long long)numel);
    if (n) read_data(m->file, d, out);
  });
}

mlf_status mlf_mat_read_f64(mlf_mat m, const char* name, double* out, int64_t numel) {
  mlf::clear_error();
  if (!m || (!out && numel)) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  return guard([&] {
    const Dataset d = lookup(m, name);
    check_readable(d);
    const int64_t n = d.empty ? 0 : (int64_t)d.numel();
    if (n != numel)
      fail(MLF_ERR_INVALID_ARGUMENT, "variable '%s' has %lld elements, buffer holds %lld", name,
           (long long)n, (long long)numel);
    if (n) read_data(m->file, d, out);
  });
}

}  // extern "C"

// ----------------------------------------------------------- PSF loader
extern "C" mlf_status mlf_psf_load_mat(mlf_context ctx, const char* path, mlf_psf* out,
                                       mlf_psf_file_info* meta) {
  mlf::clear_error();
  if (!ctx || !path || !out) return mlf::set_error(MLF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  mlf_mat m = nullptr;
  MLF_CHECK(mlf_mat_open(path, &m));
  std::unique_ptr<mlf_mat_s> hold(m);
  mlf_psf_file_info fi;
  std::memset(&fi, 0, sizeof(fi));
  std::vector<float> H, Ht;
  mlf_status st = guard([&] {
    mlf_mat_var vh, vt;
    info_of(m->file, lookup(m, "H"), &vh);
    info_of(m->file, lookup(m, "Ht"), &vt);
    if (!vh.numeric || !vt.numeric) fail(MLF_ERR_UNSUPPORTED, "H and Ht must be real numeric arrays");
    if ((vh.ndim != 4 && vh.ndim != 5) || vt.ndim != vh.ndim)
      fail(MLF_ERR_INVALID_ARGUMENT, "H and Ht must be [S,S,N,N,Z] (rank %d and %d)", vh.ndim, vt.ndim);
    for (int j = 0; j < vh.ndim; ++j)
      if (vh.dims[j] != vt.dims[j]) fail(MLF_ERR_INVALID_ARGUMENT, "H and Ht differ in size");
    if (vh.dims[0] != vh.dims[1] || vh.dims[2] != vh.dims[3])
      fail(MLF_ERR_INVALID_ARGUMENT, "H must be [S,S,N,N,Z]");
    fi.S = vh.dims[0];
    fi.N = vh.dims[2];
    fi.Z = vh.ndim == 5 ? vh.dims[4] : 1;
    for (const auto& v : m->vars)
      if (v.first == "Nnum") {
        double nn = 0;
        const Dataset d = lookup(m, "Nnum");
        check_readable(d);
        if (d.numel() != 1) fail(MLF_ERR_INVALID_ARGUMENT, "Nnum is not a scalar");
        read_data(m->file, d, &nn);
        if (nn != (double)fi.N)
          fail(MLF_ERR_INVALID_ARGUMENT, "Nnum = %g but H has %lld phases per axis", nn, (long long)fi.N);
      }
    fi.x_pixel_um = fi.z_pixel_um = std::nan("");
    auto step = [&](const char* name, double* um) {
      for (const auto& v : m->vars)
        if (v.first == name) {
          const Dataset d = lookup(m, name);
          if (d.is_group || d.numel() < 2 || (d.type.cls != 0 && d.type.cls != 1)) return;
          std::vector<double> a((size_t)d.numel());
          read_data(m->file, d, a.data());
          *um = std::fabs(a[1] - a[0]) * 1e6;  // mesolf.m:231-232
        }
    };
    step("x1objspace", &fi.x_pixel_um);
    step("x3objspace", &fi.z_pixel_um);
    H.resize((size_t)vh.numel);
    Ht.resize((size_t)vt.numel);
    read_data(m->file, lookup(m, "H"), H.data());
    read_data(m->file, lookup(m, "Ht"), Ht.data());
  });
  if (st != MLF_OK) return st;
  hold.reset();
  st = mlf_psf_create_dense(ctx, H.data(), Ht.data(), fi.S, fi.N, fi.Z, out);
  if (st == MLF_OK && meta) *meta = fi;
  return st;
}
