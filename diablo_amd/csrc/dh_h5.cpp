// Host-side reader for the subset of HDF5 the reference's grid files use
// (the /grids/y dataset of ReadGridHDF5, hdf5.f:548-710, written by
// pre_process/create_grid_h5.m): a version 0 or 1 superblock with 8-byte
// offsets and lengths, groups stored as symbol tables (version 1 B-trees of
// symbol nodes and a local heap), version 1 object headers (with
// continuation blocks), and datasets of little-endian IEEE float64 in
// contiguous or compact storage without filters.  Anything else is
// reported as unsupported.  There is no HDF5 library in this build; the
// format follows the published HDF5 file format specification (version 2.0,
// sections II-IV).  Every read is bounds-checked against the file.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/diablo_hip.h"

namespace {

struct H5Error {
  std::string msg;
};

struct File {
  std::vector<uint8_t> b;
  uint64_t base = 0;

  void need(uint64_t a, uint64_t n) const {
    if (a > b.size() || n > b.size() - a) throw H5Error{"truncated file or address out of range"};
  }
  uint64_t u(uint64_t a, int n) const {  // little-endian unsigned of n bytes
    need(a, n);
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; --i) v = (v << 8) | b[a + i];
    return v;
  }
  uint64_t addr(uint64_t a) const { return base + u(a, 8); }
  bool sig(uint64_t a, const char *s) const {
    need(a, 4);
    return std::memcmp(&b[a], s, 4) == 0;
  }
};

struct Msg {
  int type;
  uint64_t at, size;
};

// Messages of a version 1 object header, following continuation blocks.
std::vector<Msg> object_messages(const File &f, uint64_t oh) {
  if (f.u(oh, 1) != 1) throw H5Error{"object header version is not 1"};
  const uint64_t nmsg = f.u(oh + 2, 2), hsize = f.u(oh + 8, 4);
  std::vector<Msg> out;
  std::vector<std::pair<uint64_t, uint64_t>> blocks{{oh + 16, hsize}};
  for (size_t k = 0; k < blocks.size() && out.size() < nmsg; ++k) {
    uint64_t p = blocks[k].first;
    const uint64_t end = p + blocks[k].second;
    f.need(p, blocks[k].second);
    while (p + 8 <= end && out.size() < nmsg) {
      const int type = (int)f.u(p, 2);
      const uint64_t size = f.u(p + 2, 2);
      if (p + 8 + size > end) throw H5Error{"object header message overruns its block"};
      if (type == 0x10) {  // continuation: offset, length
        blocks.push_back({f.addr(p + 8), f.u(p + 16, 8)});
        if (blocks.size() > 64) throw H5Error{"too many header continuation blocks"};
      }
      out.push_back({type, p + 8, size});
      p += 8 + size;
    }
  }
  return out;
}

std::string heap_name(const File &f, uint64_t heap, uint64_t off) {
  if (!f.sig(heap, "HEAP")) throw H5Error{"local heap signature missing"};
  const uint64_t dsize = f.u(heap + 8, 8), data = f.addr(heap + 24);
  if (off >= dsize) throw H5Error{"link name offset outside the local heap"};
  f.need(data, dsize);
  std::string s;
  for (uint64_t i = data + off; i < data + dsize && f.b[i]; ++i) s.push_back((char)f.b[i]);
  return s;
}

// Object header address of member `name` of a symbol-table group.
bool find_in_btree(const File &f, uint64_t node, uint64_t heap, const std::string &name, uint64_t &oh, int depth,
                   int &visits) {
  if (depth > 32) throw H5Error{"group B-tree too deep"};
  if (++visits > 100000) throw H5Error{"group B-tree has too many nodes (a cycle?)"};
  if (!f.sig(node, "TREE")) throw H5Error{"B-tree node signature missing"};
  if (f.u(node + 4, 1) != 0) throw H5Error{"B-tree is not a group node"};
  const int level = (int)f.u(node + 5, 1), nused = (int)f.u(node + 6, 2);
  uint64_t p = node + 8 + 16;  // after the sibling addresses
  for (int i = 0; i < nused; ++i) {
    p += 8;  // key: heap offset (size of lengths)
    const uint64_t child = f.addr(p);
    p += 8;
    if (level > 0) {
      if (find_in_btree(f, child, heap, name, oh, depth + 1, visits)) return true;
      continue;
    }
    if (!f.sig(child, "SNOD")) throw H5Error{"symbol table node signature missing"};
    const int nsym = (int)f.u(child + 6, 2);
    for (int k = 0; k < nsym; ++k) {
      const uint64_t e = child + 8 + 40ull * k;
      if (heap_name(f, heap, f.u(e, 8)) == name) {
        oh = f.addr(e + 8);
        return true;
      }
    }
  }
  return false;
}

void group_tables(const File &f, uint64_t oh, uint64_t &btree, uint64_t &heap) {
  for (const Msg &m : object_messages(f, oh))
    if (m.type == 0x11) {
      btree = f.addr(m.at);
      heap = f.addr(m.at + 8);
      return;
    }
  throw H5Error{"object is not a symbol-table group"};
}

// Reads dataset `path` ("/grids/y") as float64 into out (at most cap values);
// *nout is its size, also when out is too small.
int64_t read_f64(File &f, const std::string &path, double *out, int64_t cap, int64_t *nout) {
  if (f.b.size() < 96 || std::memcmp(f.b.data(), "\x89HDF\r\n\x1a\n", 8) != 0) throw H5Error{"not an HDF5 file"};
  const int sbv = (int)f.u(8, 1);
  if (sbv != 0 && sbv != 1) throw H5Error{"superblock version is not 0 or 1"};
  if (f.u(13, 1) != 8 || f.u(14, 1) != 8) throw H5Error{"offsets and lengths must be 8 bytes"};
  const uint64_t p0 = (sbv == 0) ? 24 : 28;
  f.base = f.u(p0, 8);
  const uint64_t root_entry = p0 + 32;
  uint64_t oh = f.addr(root_entry + 8);
  // walk the path
  size_t pos = 0;
  std::vector<std::string> parts;
  while (pos < path.size()) {
    size_t q = path.find('/', pos);
    if (q == std::string::npos) q = path.size();
    if (q > pos) parts.push_back(path.substr(pos, q - pos));
    pos = q + 1;
  }
  if (parts.empty()) throw H5Error{"empty dataset path"};
  for (const std::string &name : parts) {
    uint64_t btree = 0, heap = 0;
    group_tables(f, oh, btree, heap);
    uint64_t child = 0;
    int visits = 0;
    if (!find_in_btree(f, btree, heap, name, child, 0, visits)) throw H5Error{"no member '" + name + "' in " + path};
    oh = child;
  }
  // dataset messages
  int64_t n = -1;
  bool f64 = false;
  uint64_t data_at = 0, data_size = 0;
  bool have_layout = false;
  for (const Msg &m : object_messages(f, oh)) {
    if (m.type == 0x01) {  // dataspace
      const int v = (int)f.u(m.at, 1), rank = (int)f.u(m.at + 1, 1);
      const uint64_t dims = (v == 1) ? m.at + 8 : m.at + 4;
      if (v != 1 && v != 2) throw H5Error{"dataspace version not supported"};
      n = 1;
      for (int r = 0; r < rank; ++r) {
        const uint64_t d = f.u(dims + 8ull * r, 8);
        // n <= 2^40 here, so the product bound is tested without forming it
        if (d > (1ull << 40) || (d != 0 && (uint64_t)n > (1ull << 40) / d)) throw H5Error{"dataspace too large"};
        n *= (int64_t)d;
      }
    } else if (m.type == 0x03) {  // datatype
      const int cls = (int)(f.u(m.at, 1) & 0x0f);
      const int bits0 = (int)f.u(m.at + 1, 1);
      const uint64_t size = f.u(m.at + 4, 4);
      const uint64_t prec = f.u(m.at + 10, 2);
      f64 = (cls == 1 && size == 8 && (bits0 & 0x41) == 0 && prec == 64);
    } else if (m.type == 0x0b) {
      throw H5Error{"filtered (compressed) datasets are not supported"};
    } else if (m.type == 0x08) {  // layout
      const int v = (int)f.u(m.at, 1);
      if (v != 3) throw H5Error{"data layout version not supported"};
      const int cls = (int)f.u(m.at + 1, 1);
      if (cls == 1) {
        data_at = f.addr(m.at + 2);
        data_size = f.u(m.at + 10, 8);
      } else if (cls == 0) {
        data_size = f.u(m.at + 2, 2);
        data_at = m.at + 4;
      } else {
        throw H5Error{"chunked datasets are not supported"};
      }
      have_layout = true;
    }
  }
  if (n < 0 || !have_layout) throw H5Error{path + " is not a dataset"};
  if (!f64) throw H5Error{path + " is not little-endian IEEE float64"};
  if (data_size != (uint64_t)n * 8) throw H5Error{"storage size does not match the dataspace"};
  *nout = n;
  if (n > cap) throw H5Error{"output buffer too small"};
  f.need(data_at, data_size);
  if (n > 0) std::memcpy(out, &f.b[data_at], (size_t)data_size);
  return n;
}

void put_err(char *err, int32_t len, const std::string &s) {
  if (err && len > 0) {
    std::snprintf(err, (size_t)len, "%s", s.c_str());
  }
}

}  // namespace

extern "C" int dh_h5_read_f64(const char *path, const char *dataset, double *out, int64_t cap, int64_t *n, char *err,
                              int32_t errlen) {
  put_err(err, errlen, "");
  if (!path || !dataset || !n || (cap > 0 && !out)) return DH_EARG;
  File f;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    put_err(err, errlen, std::string("cannot open ") + path);
    return DH_EARG;
  }
  std::fseek(fp, 0, SEEK_END);
  const long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (sz < 0 || sz > (1l << 34)) {
    std::fclose(fp);
    put_err(err, errlen, "file size not usable");
    return DH_EARG;
  }
  f.b.resize((size_t)sz);
  const size_t got = sz ? std::fread(f.b.data(), 1, (size_t)sz, fp) : 0;
  std::fclose(fp);
  if (got != (size_t)sz) {
    put_err(err, errlen, "short read");
    return DH_EARG;
  }
  try {
    *n = 0;
    *n = read_f64(f, dataset, out, cap, n);
  } catch (const H5Error &e) {
    put_err(err, errlen, e.msg);
    return DH_ELAYOUT;
  }
  return DH_OK;
}
