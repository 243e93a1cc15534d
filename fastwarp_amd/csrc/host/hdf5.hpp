// A small HDF5 container for ITK transform files (Composite.h5, InverseComposite.h5), written
// from the public HDF5 File Format Specification (version 3.0); the reference's own subset is
// crates/fastwarp-reg/src/transform/hdf5.rs.
//
// Writing: superblock version 2 (8-byte offsets and lengths), version-2 object headers, groups
// with compact link storage, rank-1 contiguous little-endian f64 datasets, and variable-length
// ASCII strings held in one global heap collection. A float field is widened to f64 as its bytes
// are written, so a 1 mm field is never held twice in host memory.
//
// Reading: superblock versions 2 and 3, version-2 object headers (timestamps, attribute phase
// values and message creation order allowed) with continuation blocks, every metadata checksum
// verified; compact groups of hard links; datasets of rank 0 or 1 with compact or contiguous
// layout; IEEE floats of 4 or 8 bytes in either byte order, fixed-length strings and
// variable-length strings. Anything else (version 0/1 superblocks, symbol-table or dense groups,
// chunked or filtered datasets) is refused with a message naming it.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <utility>
#include <vector>

#include <unistd.h>

#include "common.hpp"

namespace fw::h5 {

constexpr const char* kWhat = "HDF5";
constexpr uint8_t kSignature[8] = {0x89, 'H', 'D', 'F', 0x0d, 0x0a, 0x1a, 0x0a};

// Bob Jenkins' lookup3 `hashlittle` (2006, public domain) with initial value 0: the checksum of
// every piece of version-2 metadata. Known answers: "" -> 0xdeadbeef, "Four score and seven years
// ago" -> 0x17770551.
inline uint32_t lookup3(const uint8_t* k, size_t length) {
  const auto rot = [](uint32_t x, int r) { return (x << r) | (x >> (32 - r)); };
  uint32_t a = 0xdeadbeefu + uint32_t(length), b = a, c = a;
  const auto word = [&](size_t at, size_t n) {
    uint32_t w = 0;
    for (size_t i = 0; i < n; ++i) w |= uint32_t(k[at + i]) << (8 * i);
    return w;
  };
  size_t at = 0;
  while (length > 12) {
    a += word(at, 4);
    b += word(at + 4, 4);
    c += word(at + 8, 4);
    a -= c; a ^= rot(c, 4);  c += b;
    b -= a; b ^= rot(a, 6);  a += c;
    c -= b; c ^= rot(b, 8);  b += a;
    a -= c; a ^= rot(c, 16); c += b;
    b -= a; b ^= rot(a, 19); a += c;
    c -= b; c ^= rot(b, 4);  b += a;
    at += 12;
    length -= 12;
  }
  if (length == 0) return c;
  // the last 1 to 12 bytes, read as zero-padded little-endian words
  a += word(at, std::min<size_t>(length, 4));
  if (length > 4) b += word(at + 4, std::min<size_t>(length - 4, 4));
  if (length > 8) c += word(at + 8, length - 8);
  c ^= b; c -= rot(b, 14);
  a ^= c; a -= rot(c, 11);
  b ^= a; b -= rot(a, 25);
  c ^= b; c -= rot(b, 16);
  a ^= c; a -= rot(c, 4);
  b ^= a; b -= rot(a, 14);
  c ^= b; c -= rot(b, 24);
  return c;
}

inline uint32_t lookup3(const std::vector<uint8_t>& v) { return lookup3(v.data(), v.size()); }

// ------------------------------------------------------------------------------- writing --

// A member of a group to write: a group, an f64 dataset (values held, or f32 values widened when
// written), or a one-element variable-length string.
struct Node {
  enum class Kind { Group, Doubles, Floats, Text } kind = Kind::Group;
  std::string name;
  std::vector<Node> members;
  std::vector<double> doubles;
  const float* floats = nullptr;
  size_t count = 0;
  std::string text;

  static Node group(std::string name, std::vector<Node> members) {
    Node n;
    n.kind = Kind::Group;
    n.name = std::move(name);
    n.members = std::move(members);
    return n;
  }
  static Node values(std::string name, std::vector<double> v) {
    Node n;
    n.kind = Kind::Doubles;
    n.name = std::move(name);
    n.doubles = std::move(v);
    return n;
  }
  // `v` must outlive the write.
  static Node widened(std::string name, const float* v, size_t count) {
    Node n;
    n.kind = Kind::Floats;
    n.name = std::move(name);
    n.floats = v;
    n.count = count;
    return n;
  }
  static Node string(std::string name, std::string text) {
    Node n;
    n.kind = Kind::Text;
    n.name = std::move(name);
    n.text = std::move(text);
    return n;
  }
};

namespace detail {

inline void put(std::vector<uint8_t>& out, uint64_t v, size_t bytes) {
  for (size_t i = 0; i < bytes; ++i) out.push_back(uint8_t(v >> (8 * i)));
}

inline void pad8(std::vector<uint8_t>& out) {
  while (out.size() % 8) out.push_back(0);
}

struct Message {
  uint8_t type, flags;
  std::vector<uint8_t> body;
};

// A version-2 object header in one chunk, without timestamps; the chunk-size field is as wide as
// the chunk needs.
inline std::vector<uint8_t> object_header(const std::vector<Message>& msgs) {
  size_t chunk = 0;
  for (const Message& m : msgs) {
    if (m.body.size() > 0xffff) invalid(kWhat, "object header message above 64 KiB");
    chunk += 4 + m.body.size();
  }
  const int width_code = chunk <= 0xff ? 0 : (chunk <= 0xffff ? 1 : 2);
  std::vector<uint8_t> out = {'O', 'H', 'D', 'R', 2, uint8_t(width_code)};
  put(out, chunk, size_t(1) << width_code);
  for (const Message& m : msgs) {
    out.push_back(m.type);
    put(out, m.body.size(), 2);
    out.push_back(m.flags);
    out.insert(out.end(), m.body.begin(), m.body.end());
  }
  put(out, lookup3(out), 4);
  return out;
}

// Dataspace (version 2): rank 1 of n elements, maximum size n.
inline std::vector<uint8_t> dataspace_1d(uint64_t n) {
  std::vector<uint8_t> b = {2, 1, 1, 1};
  put(b, n, 8);
  put(b, n, 8);
  return b;
}

// Datatype: IEEE 754 binary64, little-endian (class 1, version 1: sign at bit 63, exponent 52..62
// with bias 1023, mantissa 0..51, implied leading one).
inline std::vector<uint8_t> f64_type() {
  std::vector<uint8_t> b = {0x11, 0x20, 63, 0};
  put(b, 8, 4);     // size
  put(b, 0, 2);     // bit offset
  put(b, 64, 2);    // precision
  b.push_back(52);  // exponent location
  b.push_back(11);  // exponent size
  b.push_back(0);   // mantissa location
  b.push_back(52);  // mantissa size
  put(b, 1023, 4);  // exponent bias
  return b;
}

// Datatype: variable-length string (class 9, version 1, type 1 = string, null-terminated padding,
// ASCII) whose base type is an unsigned byte (class 0, version 1, 8 bits).
inline std::vector<uint8_t> vlen_string_type() {
  std::vector<uint8_t> b = {0x19, 0x01, 0, 0};
  put(b, 16, 4);  // element: length (4), collection address (8), object index (4)
  b.insert(b.end(), {0x10, 0, 0, 0});
  put(b, 1, 4);
  put(b, 0, 2);
  put(b, 8, 2);
  return b;
}

// Data layout (version 3), contiguous.
inline std::vector<uint8_t> contiguous(uint64_t addr, uint64_t size) {
  std::vector<uint8_t> b = {3, 1};
  put(b, addr, 8);
  put(b, size, 8);
  return b;
}

// The file as a list of pieces written in order: metadata bytes, or f32 values written as f64.
class Layout {
 public:
  uint64_t size() const { return size_; }
  void bytes(std::vector<uint8_t> b) {
    size_ += b.size();
    pieces_.push_back({std::move(b), nullptr, 0});
  }
  void widened(const float* v, size_t n) {
    size_ += 8 * uint64_t(n);
    pieces_.push_back({{}, v, n});
  }
  void align8() {
    if (size_ % 8) bytes(std::vector<uint8_t>(size_t(8 - size_ % 8), 0));
  }
  std::vector<uint8_t>& first() { return pieces_.front().bytes; }

  void write(std::ostream& out) const {
    std::vector<uint8_t> block;
    for (const Piece& p : pieces_) {
      if (!p.floats) {
        out.write(reinterpret_cast<const char*>(p.bytes.data()), std::streamsize(p.bytes.size()));
        continue;
      }
      constexpr size_t kBlock = size_t(1) << 16;
      for (size_t s = 0; s < p.n; s += kBlock) {
        const size_t e = std::min(p.n, s + kBlock);
        block.clear();
        for (size_t i = s; i < e; ++i) {
          const double d = double(p.floats[i]);
          uint64_t bits;
          std::memcpy(&bits, &d, 8);
          put(block, bits, 8);
        }
        out.write(reinterpret_cast<const char*>(block.data()), std::streamsize(block.size()));
      }
    }
  }

 private:
  struct Piece {
    std::vector<uint8_t> bytes;
    const float* floats;
    size_t n;
  };
  std::vector<Piece> pieces_;
  uint64_t size_ = 0;
};

class Writer {
 public:
  explicit Writer(uint64_t heap) : heap_(heap) {}
  Layout layout;

  // Children before their group, so every link names an address already placed.
  uint64_t place(const Node& node) {
    switch (node.kind) {
      case Node::Kind::Group: return group(node.members);
      case Node::Kind::Doubles: {
        layout.align8();
        const uint64_t data = layout.size();
        std::vector<uint8_t> raw;
        raw.reserve(8 * node.doubles.size());
        for (double d : node.doubles) {
          uint64_t bits;
          std::memcpy(&bits, &d, 8);
          put(raw, bits, 8);
        }
        layout.bytes(std::move(raw));
        return dataset(dataspace_1d(node.doubles.size()), f64_type(), data, 8 * node.doubles.size(), 0x0a);
      }
      case Node::Kind::Floats: {
        layout.align8();
        const uint64_t data = layout.size();
        layout.widened(node.floats, node.count);
        return dataset(dataspace_1d(node.count), f64_type(), data, 8 * uint64_t(node.count), 0x0a);
      }
      case Node::Kind::Text: {
        // one element: its length, the heap collection and the string's object index
        layout.align8();
        const uint64_t data = layout.size();
        std::vector<uint8_t> raw;
        put(raw, node.text.size(), 4);
        put(raw, heap_, 8);
        put(raw, ++next_string_, 4);
        layout.bytes(std::move(raw));
        return dataset(dataspace_1d(1), vlen_string_type(), data, 16, 0x02);
      }
    }
    return 0;
  }

  uint64_t group(const std::vector<Node>& members) {
    std::vector<uint64_t> addrs;
    for (const Node& m : members) addrs.push_back(place(m));
    std::vector<Message> msgs;
    // link info (version 0, no creation order): no fractal heap and no name index, the links are
    // compact; group info (version 0, defaults)
    std::vector<uint8_t> info = {0, 0};
    put(info, ~uint64_t(0), 8);
    put(info, ~uint64_t(0), 8);
    msgs.push_back({0x02, 0, std::move(info)});
    msgs.push_back({0x0a, 1, {0, 0}});
    for (size_t i = 0; i < members.size(); ++i) {
      const std::string& name = members[i].name;
      if (name.empty() || name.size() > 0xffff) invalid(kWhat, "member names take 1 to 65535 bytes");
      // link (version 1): hard link, name length in one or two bytes
      std::vector<uint8_t> link = {1, uint8_t(name.size() <= 0xff ? 0 : 1)};
      put(link, name.size(), name.size() <= 0xff ? 1 : 2);
      link.insert(link.end(), name.begin(), name.end());
      put(link, addrs[i], 8);
      msgs.push_back({0x06, 0, std::move(link)});
    }
    layout.align8();
    const uint64_t at = layout.size();
    layout.bytes(object_header(msgs));
    return at;
  }

 private:
  // The messages and flags libhdf5 writes for a contiguous dataset: dataspace, datatype (constant),
  // fill value (version 3, constant: space allocated late; a float's fill value written if set,
  // a string's when allocated) and layout.
  uint64_t dataset(std::vector<uint8_t> space, std::vector<uint8_t> type, uint64_t data, uint64_t size, uint8_t fill) {
    layout.align8();
    const uint64_t at = layout.size();
    layout.bytes(object_header({{0x01, 0, std::move(space)},
                                {0x03, 1, std::move(type)},
                                {0x05, 1, {3, fill}},
                                {0x08, 0, contiguous(data, size)}}));
    return at;
  }

  uint64_t heap_;
  uint32_t next_string_ = 0;
};

inline void collect_strings(const std::vector<Node>& members, std::vector<const std::string*>& out) {
  for (const Node& m : members) {
    if (m.kind == Node::Kind::Group) collect_strings(m.members, out);
    if (m.kind == Node::Kind::Text) out.push_back(&m.text);
  }
}

}  // namespace detail

// The root group's members to `out`. The strings go to one global heap collection after the
// superblock, numbered in the order the writer meets them (depth first), as it places the
// datasets that refer to them.
inline void write(std::ostream& out, const std::vector<Node>& root) {
  using detail::put;
  std::vector<const std::string*> strings;
  detail::collect_strings(root, strings);
  std::vector<uint8_t> head(48, 0);  // the superblock, filled in last
  uint64_t heap = 0;
  if (!strings.empty()) {
    heap = head.size();
    std::vector<uint8_t> objects;
    for (size_t i = 0; i < strings.size(); ++i) {
      if (i + 1 > 0xffff) invalid(kWhat, "more than 65535 strings");
      put(objects, i + 1, 2);  // object index
      put(objects, 1, 2);      // reference count
      put(objects, 0, 4);
      put(objects, strings[i]->size(), 8);
      objects.insert(objects.end(), strings[i]->begin(), strings[i]->end());
      detail::pad8(objects);
    }
    const size_t used = 16 + objects.size();
    const size_t total = std::max<size_t>(used, 4096);  // collections take at least 4 KiB
    std::vector<uint8_t> gcol = {'G', 'C', 'O', 'L', 1, 0, 0, 0};
    put(gcol, total, 8);
    gcol.insert(gcol.end(), objects.begin(), objects.end());
    if (total > used) {
      // object 0 is the free space to the end of the collection
      put(gcol, 0, 8);
      put(gcol, total - used, 8);
      gcol.resize(total, 0);
    }
    head.insert(head.end(), gcol.begin(), gcol.end());
  }
  detail::Writer w(heap);
  w.layout.bytes(std::move(head));
  const uint64_t root_addr = w.group(root);
  w.layout.align8();
  std::vector<uint8_t> sb(kSignature, kSignature + 8);
  sb.insert(sb.end(), {2, 8, 8, 0});  // version 2, 8-byte offsets and lengths, no flags
  put(sb, 0, 8);                      // base address
  put(sb, ~uint64_t(0), 8);           // no superblock extension
  put(sb, w.layout.size(), 8);        // end of file
  put(sb, root_addr, 8);
  put(sb, lookup3(sb), 4);
  std::copy(sb.begin(), sb.end(), w.layout.first().begin());
  w.layout.write(out);
}

// Written to a temporary file beside `path` and renamed over it once complete, so a failed write
// leaves no partial file (the reference's fs_write.rs `write_complete`); the temporary file is
// removed on failure.
inline void write_file(const std::string& path, const std::vector<Node>& root) {
  static std::atomic<unsigned> serial{0};
  const std::string tmp = path + ".partial-" + std::to_string(::getpid()) + "-" + std::to_string(serial++);
  try {
    {
      std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
      if (!f) invalid("io", "cannot open " + tmp + " for writing");
      write(f, root);
      f.flush();
      if (!f) invalid("io", "cannot write " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) invalid("io", "cannot rename " + tmp + " to " + path);
  } catch (...) {
    std::remove(tmp.c_str());
    throw;
  }
}

// ------------------------------------------------------------------------------- reading --

class File {
 public:
  explicit File(std::vector<uint8_t> bytes) : d_(std::move(bytes)) {
    if (d_.size() < 12 || std::memcmp(d_.data(), kSignature, 8) != 0) invalid(kWhat, "not an HDF5 file (bad signature)");
    const uint8_t version = d_[8];
    if (version != 2 && version != 3)
      invalid(kWhat, "superblock version " + std::to_string(version) + " (only versions 2 and 3 are read)");
    off_ = d_[9];
    len_ = d_[10];
    if (off_ < 1 || off_ > 8 || len_ < 1 || len_ > 8) invalid(kWhat, "bad offset or length size");
    const size_t end = 12 + 4 * size_t(off_);
    if (u32(end) != lookup3(d_.data(), end)) invalid(kWhat, "superblock checksum mismatch");
    root_ = uint(12 + 3 * size_t(off_), off_);
  }

  static File open(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) invalid("io", "cannot open " + path);
    return File(std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>()));
  }

  // Member names of the group at `path` ("" is the root) in storage order.
  std::vector<std::string> members(const std::string& path) const {
    std::vector<std::string> out;
    for (const auto& l : links(resolve(path))) out.push_back(l.first);
    return out;
  }

  // The header messages of the object at `path` in storage order (type and body; continuation
  // blocks followed), for checks of the format itself.
  std::vector<std::pair<uint8_t, std::vector<uint8_t>>> header_messages(const std::string& path) const {
    std::vector<std::pair<uint8_t, std::vector<uint8_t>>> out;
    for (const Msg& m : messages(resolve(path)))
      out.emplace_back(m.type, std::vector<uint8_t>(d_.begin() + std::ptrdiff_t(m.at), d_.begin() + std::ptrdiff_t(m.at + m.size)));
    return out;
  }

  // Whether `path` names a group (its header holds link or group information).
  bool is_group(const std::string& path) const {
    for (const Msg& m : messages(resolve(path)))
      if (m.type == 0x02 || m.type == 0x06 || m.type == 0x0a || m.type == 0x11) return true;
    return false;
  }

  bool exists(const std::string& path) const {
    try {
      resolve(path);
      return true;
    } catch (const std::exception&) {
      return false;
    }
  }

  // A float dataset as f64 (f32 values widened).
  std::vector<double> read_doubles(const std::string& path) const {
    const Dataset ds = dataset(resolve(path));
    if (ds.type_class != 1) invalid(kWhat, "`" + path + "` is not a float dataset");
    if (ds.elem != 4 && ds.elem != 8) invalid(kWhat, "float of " + std::to_string(ds.elem) + " bytes");
    std::vector<double> out(ds.count);
    for (uint64_t i = 0; i < ds.count; ++i) {
      const size_t at = ds.data + size_t(i) * ds.elem;
      uint64_t bits = 0;
      for (size_t b = 0; b < ds.elem; ++b) bits |= uint64_t(d_[at + b]) << (8 * (ds.big_endian ? ds.elem - 1 - b : b));
      if (ds.elem == 8) {
        std::memcpy(&out[i], &bits, 8);
      } else {
        const uint32_t b32 = uint32_t(bits);
        float f;
        std::memcpy(&f, &b32, 4);
        out[i] = double(f);
      }
    }
    return out;
  }

  // A string dataset of one element (variable-length or fixed-length), without trailing NULs.
  std::string read_string(const std::string& path) const {
    const Dataset ds = dataset(resolve(path));
    if (ds.count != 1) invalid(kWhat, "`" + path + "` is not a single string");
    std::string s;
    if (ds.type_class == 3) {
      s.assign(reinterpret_cast<const char*>(&d_[ds.data]), ds.elem);
    } else if (ds.type_class == 9 && ds.vlen_string) {
      const uint64_t n = u32(ds.data);
      const uint64_t collection = uint(ds.data + 4, off_);
      const uint32_t index = u32(ds.data + 4 + off_);
      s = heap_object(collection, index);
      if (s.size() < n) invalid(kWhat, "variable-length string longer than its heap object");
      s.resize(size_t(n));
    } else {
      invalid(kWhat, "`" + path + "` is not a string dataset");
    }
    while (!s.empty() && s.back() == '\0') s.pop_back();
    return s;
  }

 private:
  struct Msg {
    uint8_t type;
    size_t at, size;
  };
  struct Dataset {
    uint64_t count = 0;
    int type_class = -1;
    size_t elem = 0;
    bool big_endian = false, vlen_string = false;
    size_t data = 0;  // file offset of the first element
  };

  void need(uint64_t at, uint64_t n) const {
    if (at > d_.size() || n > d_.size() - at) invalid(kWhat, "read past the end of the file");
  }
  uint64_t uint(uint64_t at, size_t n) const {
    need(at, n);
    uint64_t v = 0;
    for (size_t i = 0; i < n; ++i) v |= uint64_t(d_[size_t(at) + i]) << (8 * i);
    return v;
  }
  uint32_t u32(uint64_t at) const { return uint32_t(uint(at, 4)); }
  bool undefined(uint64_t addr) const { return off_ == 8 ? addr == ~uint64_t(0) : addr == (uint64_t(1) << (8 * off_)) - 1; }

  // The messages of the object header at `addr`, continuation blocks followed.
  std::vector<Msg> messages(uint64_t addr) const {
    need(addr, 6);
    if (std::memcmp(&d_[size_t(addr)], "OHDR", 4) != 0)
      invalid(kWhat, "no version-2 object header at " + std::to_string(addr) + " (version-1 headers are not read)");
    if (d_[size_t(addr) + 4] != 2) invalid(kWhat, "object header version " + std::to_string(d_[size_t(addr) + 4]));
    const uint8_t flags = d_[size_t(addr) + 5];
    uint64_t pos = addr + 6;
    if (flags & 0x20) pos += 16;  // access, modification, change and birth times
    if (flags & 0x10) pos += 4;   // attribute phase change values
    const size_t width = size_t(1) << (flags & 3);
    const uint64_t chunk = uint(pos, width);
    pos += width;
    need(pos, chunk + 4);
    const uint64_t end = pos + chunk;
    if (u32(end) != lookup3(&d_[size_t(addr)], size_t(end - addr)))
      invalid(kWhat, "object header checksum mismatch at " + std::to_string(addr));
    std::vector<Msg> out;
    std::vector<uint64_t> seen;
    walk(pos, end, (flags & 4) != 0, out, seen, 0);
    return out;
  }

  void walk(uint64_t pos, uint64_t end, bool order, std::vector<Msg>& out, std::vector<uint64_t>& seen, int depth) const {
    if (depth > 32) invalid(kWhat, "object header continuation chain too deep");
    const uint64_t head = order ? 6 : 4;
    while (pos + head <= end) {
      const uint8_t type = d_[size_t(pos)];
      const uint64_t size = uint(pos + 1, 2);
      pos += head;
      if (size > end - pos) invalid(kWhat, "object header message overruns its chunk");
      if (type == 0x10) {
        // continuation: a further block of messages, "OCHK", messages, checksum
        const uint64_t block = uint(pos, off_), length = uint(pos + off_, len_);
        if (length < 8) invalid(kWhat, "bad continuation block length");
        if (std::find(seen.begin(), seen.end(), block) != seen.end()) invalid(kWhat, "continuation block referenced twice");
        seen.push_back(block);
        need(block, length);
        if (std::memcmp(&d_[size_t(block)], "OCHK", 4) != 0) invalid(kWhat, "continuation block without its signature");
        if (u32(block + length - 4) != lookup3(&d_[size_t(block)], size_t(length - 4)))
          invalid(kWhat, "continuation block checksum mismatch");
        walk(block + 4, block + length - 4, order, out, seen, depth + 1);
      } else if (type != 0) {
        out.push_back({type, size_t(pos), size_t(size)});
      }
      pos += size;
    }
  }

  std::vector<std::pair<std::string, uint64_t>> links(uint64_t addr) const {
    std::vector<std::pair<std::string, uint64_t>> out;
    for (const Msg& m : messages(addr)) {
      if (m.type == 0x11) invalid(kWhat, "symbol-table (version 1) groups are not read");
      if (m.type == 0x02) {
        const uint8_t flags = m.size > 1 ? d_[m.at + 1] : 0;
        const size_t heap_at = m.at + 2 + ((flags & 1) ? 8 : 0);
        if (!undefined(uint(heap_at, off_))) invalid(kWhat, "dense link storage (a fractal heap) is not read");
      }
      if (m.type != 0x06) continue;
      size_t at = m.at;
      const size_t end = m.at + m.size;
      const uint8_t flags = uint8_t(uint(at + 1, 1));
      at += 2;
      uint8_t link_type = 0;
      if (flags & 0x08) link_type = uint8_t(uint(at++, 1));
      if (flags & 0x04) at += 8;  // creation order
      if (flags & 0x10) at += 1;  // character set
      const size_t width = size_t(1) << (flags & 3);
      const uint64_t name_len = uint(at, width);
      at += width;
      if (name_len > end - std::min(end, at)) invalid(kWhat, "link name overruns its message");
      std::string name(reinterpret_cast<const char*>(&d_[at]), size_t(name_len));
      at += size_t(name_len);
      if (link_type == 0) out.emplace_back(std::move(name), uint(at, off_));  // soft and external links are skipped
    }
    return out;
  }

  uint64_t resolve(const std::string& path) const {
    uint64_t addr = root_;
    size_t s = 0;
    while (s <= path.size()) {
      size_t e = path.find('/', s);
      if (e == std::string::npos) e = path.size();
      if (e > s) {
        const std::string part = path.substr(s, e - s);
        bool found = false;
        for (const auto& l : links(addr))
          if (l.first == part) {
            addr = l.second;
            found = true;
            break;
          }
        if (!found) invalid(kWhat, "no member `" + part + "` in `" + path + "`");
      }
      s = e + 1;
    }
    return addr;
  }

  Dataset dataset(uint64_t addr) const {
    const std::vector<Msg> msgs = messages(addr);
    const auto find = [&](uint8_t t) -> const Msg* {
      for (const Msg& m : msgs)
        if (m.type == t) return &m;
      return nullptr;
    };
    const Msg* space = find(0x01);
    const Msg* type = find(0x03);
    const Msg* layout = find(0x08);
    if (!space || !type || !layout) invalid(kWhat, "object at " + std::to_string(addr) + " is not a dataset");
    if (find(0x0b)) invalid(kWhat, "filtered datasets are not read");
    Dataset ds;
    // dataspace: version 1 (dimensions at 8) or 2 (at 4; type 0 scalar, 1 simple, 2 null)
    const uint8_t sv = d_[space->at], rank = uint8_t(uint(space->at + 1, 1));
    if (sv != 1 && sv != 2) invalid(kWhat, "dataspace version " + std::to_string(sv));
    if (rank > 1) invalid(kWhat, "datasets of rank above 1 are not read");
    ds.count = 1;
    if (sv == 2 && uint(space->at + 3, 1) == 2) ds.count = 0;
    if (rank == 1) ds.count = uint(space->at + (sv == 1 ? 8 : 4), len_);
    // datatype
    ds.type_class = d_[type->at] & 0x0f;
    ds.elem = size_t(u32(type->at + 4));
    ds.big_endian = ds.type_class == 1 && (d_[type->at + 1] & 1);
    ds.vlen_string = ds.type_class == 9 && (d_[type->at + 1] & 0x0f) == 1;
    if (ds.elem == 0) invalid(kWhat, "datatype of size 0");
    // layout: compact or contiguous (versions 3 and 4)
    const uint8_t lv = d_[layout->at], lclass = uint8_t(uint(layout->at + 1, 1));
    if (lv != 3 && lv != 4) invalid(kWhat, "data layout version " + std::to_string(lv));
    const uint64_t bytes = ds.count * ds.elem;
    if (ds.count && bytes / ds.count != ds.elem) invalid(kWhat, "dataset too large");
    if (lclass == 0) {
      const uint64_t size = uint(layout->at + 2, 2);
      if (size < bytes) invalid(kWhat, "compact dataset shorter than its extent");
      if (4 + size > layout->size) invalid(kWhat, "compact data overruns its message");
      ds.data = layout->at + 4;
    } else if (lclass == 1) {
      const uint64_t at = uint(layout->at + 2, off_), size = uint(layout->at + 2 + off_, len_);
      if (bytes && (undefined(at) || size < bytes)) invalid(kWhat, "contiguous dataset without its data");
      if (bytes) need(at, bytes);
      ds.data = size_t(at);
    } else {
      invalid(kWhat, "chunked datasets are not read");
    }
    return ds;
  }

  // Object `index` of the global heap collection at `collection`.
  std::string heap_object(uint64_t collection, uint32_t index) const {
    need(collection, 8 + len_);
    if (std::memcmp(&d_[size_t(collection)], "GCOL", 4) != 0) invalid(kWhat, "global heap collection without its signature");
    const uint64_t total = uint(collection + 8, len_);
    need(collection, total);
    uint64_t at = collection + 8 + len_;
    const uint64_t end = collection + total;
    while (at + 8 + len_ <= end) {
      const uint64_t id = uint(at, 2), size = uint(at + 8, len_);
      const uint64_t data = at + 8 + len_;
      if (id == 0) break;  // free space
      if (size > end - data) invalid(kWhat, "global heap object overruns its collection");
      if (id == index) return std::string(reinterpret_cast<const char*>(&d_[size_t(data)]), size_t(size));
      at = data + (size + 7) / 8 * 8;
    }
    invalid(kWhat, "no global heap object " + std::to_string(index));
    return {};
  }

  std::vector<uint8_t> d_;
  size_t off_ = 8, len_ = 8;
  uint64_t root_ = 0;
};

}  // namespace fw::h5
