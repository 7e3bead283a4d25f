// Solution files in HDF5 format, written and read without the HDF5 library.
//
// The reference writes its restart files with libhdf5
// (channel/channel.c:889-1028, save_solution; read back by read_solution,
// channel/channel.c:791-887): datasets /data_alpha, /data_beta,
// /data_ialpha, /data_ibeta of shape [dimR][Nz][Nx/2] with the compound
// type {double real; double imaginary}, and data_Nx, data_Ny, data_Nz
// (int32 LE) and data_dt, data_Re, data_mpg (float64 LE), each of shape [1].
//
// The writer lays out a file as libhdf5's defaults do (superblock version 0,
// the root group as a symbol table: a version-1 B-tree of symbol-table
// nodes and a local heap of names; version-1 object headers; contiguous
// data).  The reader walks that structure for any file of the same family
// (superblock 0 or 1, a userblock before it, version-1 object headers with
// continuation blocks, layout messages versions 1 to 3, contiguous or
// compact data) and converts fixed-point, floating-point and two-member
// floating-point compound (complex) data to double.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/shadow_channel.h"
#include "sc_internal.h"

namespace sc {
namespace h5 {

constexpr uint64_t UNDEF = ~0ull;
constexpr int LEAF_K = 4, NODE_K = 16;           // libhdf5's default group B-tree ranks

// ------------------------------------------------------------------ writer

struct Buf {
    std::vector<uint8_t> b;
    size_t tell() const { return b.size(); }
    void u8(unsigned v) { b.push_back((uint8_t) v); }
    void u16(unsigned v) { for (int i = 0; i < 2; ++i) u8((v >> (8 * i)) & 0xff); }
    void u32(uint32_t v) { for (int i = 0; i < 4; ++i) u8((v >> (8 * i)) & 0xff); }
    void u64(uint64_t v) { for (int i = 0; i < 8; ++i) u8((unsigned) ((v >> (8 * i)) & 0xff)); }
    void bytes(const void *p, size_t n) { b.insert(b.end(), (const uint8_t *) p, (const uint8_t *) p + n); }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void align8() { zeros((8 - b.size() % 8) % 8); }
};

// datatype messages (version 1)
static void dtype_f64(Buf &m)
{
    m.u8(0x11);                 // version 1, class 1 (floating point)
    m.u8(0x20);                 // little endian, mantissa normalised (implied msb)
    m.u8(63);                   // sign bit position
    m.u8(0);
    m.u32(8);
    m.u16(0);                   // bit offset
    m.u16(64);                  // precision
    m.u8(52);                   // exponent location
    m.u8(11);                   // exponent size
    m.u8(0);                    // mantissa location
    m.u8(52);                   // mantissa size
    m.u32(1023);                // exponent bias
}

static void dtype_i32(Buf &m)
{
    m.u8(0x10);                 // version 1, class 0 (fixed point)
    m.u8(0x08);                 // little endian, signed
    m.u8(0);
    m.u8(0);
    m.u32(4);
    m.u16(0);
    m.u16(32);
}

static void dtype_complex(Buf &m)
{
    m.u8(0x16);                 // version 1, class 6 (compound)
    m.u16(2);                   // members
    m.u8(0);
    m.u32(16);
    const char *names[2] = {"real", "imaginary"};
    for (int k = 0; k < 2; ++k) {
        const size_t n = strlen(names[k]) + 1;
        m.bytes(names[k], n);
        m.zeros((8 - n % 8) % 8);
        m.u32(8 * k);           // byte offset in the element
        m.u8(0);                // dimensionality (scalar member)
        m.zeros(3);
        m.u32(0);               // dimension permutation
        m.u32(0);
        m.zeros(16);            // dimension sizes
        dtype_f64(m);
    }
}

struct DatasetSpec {
    std::string name;
    int kind;                   // 0 int32, 1 float64, 2 complex128
    std::vector<uint64_t> dims;
    const void *data;
    size_t nbytes;
};

static void message(Buf &oh, unsigned type, const Buf &body, unsigned flags)
{
    oh.u16(type);
    const size_t n = (body.b.size() + 7) / 8 * 8;
    oh.u16((unsigned) n);
    oh.u8(flags);
    oh.zeros(3);
    oh.bytes(body.b.data(), body.b.size());
    oh.zeros(n - body.b.size());
}

// object header of a dataset whose data will sit at data_addr
static Buf dataset_header(const DatasetSpec &s, uint64_t data_addr)
{
    Buf msgs;
    Buf fill;                   // fill value: allocated late, written if set, default value
    fill.u8(1);
    fill.u8(2);
    fill.u8(2);
    fill.u8(1);
    fill.u32(0);
    message(msgs, 0x0005, fill, 1);
    Buf dt;
    if (s.kind == 0) dtype_i32(dt);
    else if (s.kind == 1) dtype_f64(dt);
    else dtype_complex(dt);
    message(msgs, 0x0003, dt, 1);
    Buf sp;                     // dataspace version 1, no maximum dimensions
    sp.u8(1);
    sp.u8((unsigned) s.dims.size());
    sp.u8(0);
    sp.u8(0);
    sp.u32(0);
    for (uint64_t d : s.dims) sp.u64(d);
    message(msgs, 0x0001, sp, 0);
    Buf lay;                    // layout version 3, contiguous
    lay.u8(3);
    lay.u8(1);
    lay.u64(data_addr);
    lay.u64(s.nbytes);
    message(msgs, 0x0008, lay, 0);
    Buf oh;
    oh.u8(1);                   // version 1
    oh.u8(0);
    oh.u16(4);                  // messages
    oh.u32(1);                  // reference count
    oh.u32((uint32_t) msgs.b.size());
    oh.u32(0);                  // alignment of the messages
    oh.bytes(msgs.b.data(), msgs.b.size());
    return oh;
}

static bool write_file(const std::string &path, std::vector<DatasetSpec> sets, std::string &err)
{
    std::sort(sets.begin(), sets.end(),
              [](const DatasetSpec &a, const DatasetSpec &b) { return strcmp(a.name.c_str(), b.name.c_str()) < 0; });
    const int n = (int) sets.size();
    const int nsnod = (n + 2 * LEAF_K - 1) / (2 * LEAF_K);
    if (nsnod > 2 * NODE_K) {
        err = "h5: too many datasets for one B-tree node";
        return false;
    }
    // local heap data: "" at 0, then the names, 8-byte aligned
    Buf heap;
    heap.zeros(8);
    std::vector<uint64_t> name_off(n);
    for (int i = 0; i < n; ++i) {
        name_off[i] = heap.tell();
        heap.bytes(sets[i].name.c_str(), sets[i].name.size() + 1);
        heap.align8();
    }
    // file map: superblock 96, root object header, heap header 32 + data,
    // B-tree node, symbol-table nodes, dataset headers, data
    const uint64_t sb = 96, root_oh = sb, root_oh_size = 16 + 8 + 16;
    const uint64_t heap_hdr = root_oh + root_oh_size, heap_data = heap_hdr + 32;
    const uint64_t btree = heap_data + heap.tell();
    const uint64_t btree_size = 24 + (2 * NODE_K + 1) * 8 + 2 * NODE_K * 8;
    const uint64_t snod0 = btree + btree_size, snod_size = 8 + 2 * LEAF_K * 40;
    uint64_t at = snod0 + nsnod * snod_size;
    std::vector<uint64_t> oh_addr(n), data_addr(n);
    for (int i = 0; i < n; ++i) {
        oh_addr[i] = at;
        at += dataset_header(sets[i], 0).tell();
    }
    for (int i = 0; i < n; ++i) {
        data_addr[i] = at;
        at += (sets[i].nbytes + 7) / 8 * 8;
    }
    const uint64_t eof = at;
    Buf f;
    // superblock version 0
    const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    f.bytes(sig, 8);
    f.u8(0);                    // superblock version
    f.u8(0);                    // free-space storage version
    f.u8(0);                    // root group symbol table entry version
    f.u8(0);
    f.u8(0);                    // shared header message format version
    f.u8(8);                    // size of offsets
    f.u8(8);                    // size of lengths
    f.u8(0);
    f.u16(LEAF_K);
    f.u16(NODE_K);
    f.u32(0);                   // file consistency flags
    f.u64(0);                   // base address
    f.u64(UNDEF);               // free-space info
    f.u64(eof);
    f.u64(UNDEF);               // driver information block
    // root group symbol table entry: cached B-tree and heap addresses
    f.u64(0);
    f.u64(root_oh);
    f.u32(1);
    f.u32(0);
    f.u64(btree);
    f.u64(heap_hdr);
    // root group object header: one symbol table message
    f.u8(1);
    f.u8(0);
    f.u16(1);
    f.u32(1);
    f.u32(24);
    f.u32(0);
    f.u16(0x0011);
    f.u16(16);
    f.u8(1);
    f.zeros(3);
    f.u64(btree);
    f.u64(heap_hdr);
    // local heap (no free block: the free list holds 1, libhdf5's empty mark)
    f.bytes("HEAP", 4);
    f.u8(0);
    f.zeros(3);
    f.u64(heap.tell());
    f.u64(1);
    f.u64(heap_data);
    f.bytes(heap.b.data(), heap.tell());
    // B-tree node (group, leaf level): key i = heap offset of the last name
    // of child i-1 (key 0 = the empty name), child i = symbol-table node i
    f.bytes("TREE", 4);
    f.u8(0);
    f.u8(0);
    f.u16(nsnod);
    f.u64(UNDEF);
    f.u64(UNDEF);
    const size_t keys_at = f.tell();
    f.u64(0);
    for (int s = 0; s < nsnod; ++s) {
        f.u64(snod0 + s * snod_size);
        const int last = std::min(n, (s + 1) * 2 * LEAF_K) - 1;
        f.u64(name_off[last]);
    }
    f.zeros(btree_size - (f.tell() - keys_at) - 24);
    for (int s = 0; s < nsnod; ++s) {
        f.bytes("SNOD", 4);
        f.u8(1);
        f.u8(0);
        const int i0 = s * 2 * LEAF_K, i1 = std::min(n, i0 + 2 * LEAF_K);
        f.u16(i1 - i0);
        for (int i = i0; i < i1; ++i) {
            f.u64(name_off[i]);
            f.u64(oh_addr[i]);
            f.u32(0);
            f.u32(0);
            f.zeros(16);
        }
        f.zeros((2 * LEAF_K - (i1 - i0)) * 40);
    }
    for (int i = 0; i < n; ++i) {
        Buf oh = dataset_header(sets[i], data_addr[i]);
        f.bytes(oh.b.data(), oh.tell());
    }
    for (int i = 0; i < n; ++i) {
        f.bytes(sets[i].data, sets[i].nbytes);
        f.align8();
    }
    if (f.tell() != eof) {
        err = "h5: internal layout error";
        return false;
    }
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out || !out.write((const char *) f.b.data(), (std::streamsize) f.tell())) {
        err = "cannot write " + path;
        return false;
    }
    return true;
}

// ------------------------------------------------------------------ reader

struct Dataset {
    int kind = -1;              // 0 integer, 1 floating point, 2 complex
    std::vector<uint64_t> dims;
    std::vector<double> values; // complex as (re, im) pairs
};

class Reader {
  public:
    explicit Reader(const std::string &path) : path_(path) {}

    bool open(std::string &err)
    {
        std::ifstream in(path_, std::ios::binary);
        if (!in) return fail(err, "cannot open " + path_);
        f_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
        const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        for (uint64_t at = 0; at + 8 <= f_.size(); at = at ? 2 * at : 512)
            if (memcmp(f_.data() + at, sig, 8) == 0) {
                sb_ = at;
                return superblock(err);
            }
        return fail(err, path_ + " is not an HDF5 file");
    }

    // the root group's datasets by name
    bool datasets(std::map<std::string, uint64_t> &out, std::string &err)
    {
        return group(root_oh_, out, err, 0);
    }

    bool dataset(uint64_t oh, Dataset &ds, std::string &err)
    {
        std::vector<Msg> msgs;
        if (!messages(oh, msgs, err)) return false;
        const Msg *space = nullptr, *type = nullptr, *layout = nullptr;
        for (const Msg &m : msgs) {
            if (m.type == 0x0001) space = &m;
            else if (m.type == 0x0003) type = &m;
            else if (m.type == 0x0008) layout = &m;
            else if (m.type == 0x000B) return fail(err, "filtered (compressed) data is not supported");
        }
        if (!space || !type || !layout) return fail(err, "not a dataset");
        uint64_t count = 1;
        if (!dataspace(*space, ds.dims, err)) return false;
        for (uint64_t d : ds.dims) count *= d;
        Type t;
        if (!datatype(type->at, type->size, t, err)) return false;
        const uint8_t *raw = nullptr;
        uint64_t nraw = 0;
        if (!data(*layout, ds.dims.size(), raw, nraw, err)) return false;
        if (nraw < count * t.size) return fail(err, "dataset data shorter than its dataspace");
        ds.kind = t.cls == 6 ? 2 : t.cls;
        ds.values.resize(count * (t.cls == 6 ? 2 : 1));
        for (uint64_t e = 0; e < count; ++e) {
            const uint8_t *p = raw + e * t.size;
            if (t.cls == 6) {
                ds.values[2 * e] = scalar(t.parts[0], p + t.offs[0]);
                ds.values[2 * e + 1] = scalar(t.parts[1], p + t.offs[1]);
            } else {
                ds.values[e] = scalar(t, p);
            }
        }
        return true;
    }

  private:
    struct Msg {
        unsigned type;
        uint64_t at, size;      // message body: file offset and length
    };
    struct Type {
        int cls = -1;           // 0 fixed point, 1 floating point, 6 compound
        uint32_t size = 0;
        bool big = false, sgn = false;
        std::vector<Type> parts;     // compound: real then imaginary
        std::vector<uint32_t> offs;
    };

    std::string path_;
    std::vector<uint8_t> f_;
    uint64_t sb_ = 0, base_ = 0, root_oh_ = 0;
    int so_ = 8, sl_ = 8;

    static bool fail(std::string &err, const std::string &msg)
    {
        err = msg;
        return false;
    }
    bool have(uint64_t at, uint64_t n) const { return at <= f_.size() && n <= f_.size() - at; }
    uint64_t le(uint64_t at, int n) const
    {
        uint64_t v = 0;
        for (int i = n - 1; i >= 0; --i) v = (v << 8) | f_[at + i];
        return v;
    }
    // file offset of an address (addresses are relative to the base address)
    uint64_t addr(uint64_t a) const { return a + base_; }
    bool is_undef(uint64_t a) const { return so_ == 8 ? a == UNDEF : a == (1ull << (8 * so_)) - 1; }

    bool superblock(std::string &err)
    {
        if (!have(sb_, 24)) return fail(err, "truncated superblock");
        const int ver = f_[sb_ + 8];
        if (ver > 1) return fail(err, "superblock version " + std::to_string(ver) + " is not supported");
        so_ = f_[sb_ + 13];
        sl_ = f_[sb_ + 14];
        if ((so_ != 4 && so_ != 8) || (sl_ != 4 && sl_ != 8)) return fail(err, "unsupported offset size");
        uint64_t p = sb_ + 24 + (ver == 1 ? 4 : 0);
        if (!have(p, 4 * so_ + 2 * so_ + 24)) return fail(err, "truncated superblock");
        base_ = le(p, so_);
        // a userblock: libhdf5 stores the base address as the superblock's
        // own offset; treat any other base as given
        p += 4 * so_;           // base, free space, end of file, driver block
        root_oh_ = le(p + so_, so_);
        return true;
    }

    bool messages(uint64_t oh_addr, std::vector<Msg> &out, std::string &err)
    {
        const uint64_t oh = addr(oh_addr);
        if (!have(oh, 16)) return fail(err, "object header out of the file");
        if (f_[oh] != 1) return fail(err, "object header version " + std::to_string(f_[oh]) + " is not supported");
        const unsigned nmsg = (unsigned) le(oh + 2, 2);
        std::vector<std::pair<uint64_t, uint64_t>> blocks = {{oh + 16, le(oh + 8, 4)}};
        unsigned seen = 0;
        for (size_t b = 0; b < blocks.size() && seen < nmsg; ++b) {
            uint64_t p = blocks[b].first;
            const uint64_t end = p + blocks[b].second;
            if (!have(blocks[b].first, blocks[b].second)) return fail(err, "object header out of the file");
            while (p + 8 <= end && seen < nmsg) {
                const unsigned type = (unsigned) le(p, 2);
                const uint64_t size = le(p + 2, 2);
                const unsigned flags = f_[p + 4];
                if (p + 8 + size > end) return fail(err, "object header message overruns its block");
                if (flags & 0x02) return fail(err, "shared object header messages are not supported");
                if (type == 0x0010) // continuation: another block of messages
                    blocks.push_back({addr(le(p + 8, so_)), le(p + 8 + so_, sl_)});
                else
                    out.push_back({type, p + 8, size});
                p += 8 + size;
                ++seen;
            }
        }
        return true;
    }

    // the root group: symbol table message -> B-tree -> symbol-table nodes
    bool group(uint64_t oh_addr, std::map<std::string, uint64_t> &out, std::string &err, int depth)
    {
        std::vector<Msg> msgs;
        if (!messages(oh_addr, msgs, err)) return false;
        for (const Msg &m : msgs)
            if (m.type == 0x0011) {
                const uint64_t bt = le(m.at, so_), hp = le(m.at + so_, so_);
                const uint64_t h = addr(hp);
                if (!have(h, 8 + 2 * sl_ + so_) || memcmp(f_.data() + h, "HEAP", 4) != 0)
                    return fail(err, "bad local heap");
                const uint64_t hsize = le(h + 8, sl_), hdata = addr(le(h + 8 + 2 * sl_, so_));
                if (!have(hdata, hsize)) return fail(err, "local heap out of the file");
                return btree(bt, hdata, hsize, out, err, depth);
            }
        return fail(err, "the root group has no symbol table (newer group formats are not supported)");
    }

    bool btree(uint64_t node_addr, uint64_t hdata, uint64_t hsize,
               std::map<std::string, uint64_t> &out, std::string &err, int depth)
    {
        if (depth > 16) return fail(err, "B-tree too deep");
        const uint64_t b = addr(node_addr);
        if (!have(b, 8 + 2 * so_) || memcmp(f_.data() + b, "TREE", 4) != 0 || f_[b + 4] != 0)
            return fail(err, "bad group B-tree node");
        const int level = f_[b + 5];
        const unsigned used = (unsigned) le(b + 6, 2);
        uint64_t p = b + 8 + 2 * so_ + sl_;          // first child
        for (unsigned i = 0; i < used; ++i, p += so_ + sl_) {
            if (!have(p, so_)) return fail(err, "B-tree node out of the file");
            const uint64_t child = le(p, so_);
            if (level > 0) {
                if (!btree(child, hdata, hsize, out, err, depth + 1)) return false;
                continue;
            }
            const uint64_t s = addr(child);
            if (!have(s, 8) || memcmp(f_.data() + s, "SNOD", 4) != 0)
                return fail(err, "bad symbol table node");
            const unsigned nsym = (unsigned) le(s + 6, 2);
            const uint64_t esize = 2 * so_ + 24;
            for (unsigned k = 0; k < nsym; ++k) {
                const uint64_t e = s + 8 + k * esize;
                if (!have(e, esize)) return fail(err, "symbol table node out of the file");
                const uint64_t noff = le(e, so_);
                if (noff >= hsize) return fail(err, "name outside the local heap");
                const char *nm = (const char *) f_.data() + hdata + noff;
                out[std::string(nm, strnlen(nm, hsize - noff))] = le(e + so_, so_);
            }
        }
        return true;
    }

    bool dataspace(const Msg &m, std::vector<uint64_t> &dims, std::string &err)
    {
        const unsigned ver = f_[m.at], rank = f_[m.at + 1];
        const uint64_t p = m.at + (ver == 1 ? 8 : 4);
        if (ver < 1 || ver > 2) return fail(err, "unsupported dataspace version");
        if (p + rank * sl_ > m.at + m.size) return fail(err, "bad dataspace");
        dims.clear();
        for (unsigned i = 0; i < rank; ++i) dims.push_back(le(p + i * sl_, sl_));
        return true;
    }

    bool datatype(uint64_t at, uint64_t avail, Type &t, std::string &err)
    {
        if (avail < 8) return fail(err, "bad datatype");
        const unsigned ver = f_[at] >> 4;
        t.cls = f_[at] & 0x0f;
        const unsigned bits = (unsigned) le(at + 1, 3);
        t.size = (uint32_t) le(at + 4, 4);
        t.big = bits & 1;
        if (t.cls == 0) {
            t.sgn = (bits >> 3) & 1;
            if (t.size != 1 && t.size != 2 && t.size != 4 && t.size != 8) return fail(err, "unsupported integer size");
            return true;
        }
        if (t.cls == 1) {
            if (t.size != 4 && t.size != 8) return fail(err, "unsupported floating-point size");
            return true;
        }
        if (t.cls != 6) return fail(err, "unsupported datatype class " + std::to_string(t.cls));
        const unsigned nmem = bits & 0xffff;
        uint64_t p = at + 8;
        const uint64_t end = at + avail;
        for (unsigned k = 0; k < nmem; ++k) {
            const char *nm = (const char *) f_.data() + p;
            const size_t nl = strnlen(nm, end - p);
            if (ver == 3) p += nl + 1;
            else p += (nl + 1 + 7) / 8 * 8;
            uint32_t off;
            if (ver == 3) {
                int ob = 1;
                while (ob < 8 && (t.size >> (8 * ob)) != 0) ++ob;
                off = (uint32_t) le(p, ob);
                p += ob;
            } else {
                off = (uint32_t) le(p, 4);
                p += ver == 1 ? 4 + 28 : 4;
            }
            Type mt;
            if (!datatype(p, end - p, mt, err)) return false;
            if (mt.cls != 1) return fail(err, "only floating-point compound members are supported");
            p += 8 + 12;        // floating-point member: 8 + properties
            t.parts.push_back(mt);
            t.offs.push_back(off);
        }
        if (t.parts.size() != 2) return fail(err, "only two-member (complex) compounds are supported");
        return true;
    }

    bool data(const Msg &m, size_t rank, const uint8_t *&raw, uint64_t &n, std::string &err)
    {
        const unsigned ver = f_[m.at];
        if (ver == 3) {
            const unsigned cls = f_[m.at + 1];
            if (cls == 0) {     // compact: size, data
                n = le(m.at + 2, 2);
                raw = f_.data() + m.at + 4;
                return have(m.at + 4, n) ? true : fail(err, "compact data out of the file");
            }
            if (cls != 1) return fail(err, "chunked data is not supported");
            const uint64_t a = le(m.at + 2, so_);
            n = le(m.at + 2 + so_, sl_);
            if (is_undef(a)) return fail(err, "dataset has no data");
            raw = f_.data() + addr(a);
            return have(addr(a), n) ? true : fail(err, "data out of the file");
        }
        if (ver != 1 && ver != 2) return fail(err, "unsupported layout version");
        const unsigned nd = f_[m.at + 1], cls = f_[m.at + 2];
        uint64_t p = m.at + 8;
        uint64_t a = UNDEF;
        if (cls != 0) {
            a = le(p, so_);
            p += so_;
        }
        if (cls == 2) return fail(err, "chunked data is not supported");
        // dimension sizes (the last one is the element size in libhdf5's files)
        uint64_t total = 1;
        for (unsigned i = 0; i < nd; ++i) total *= le(p + 4 * i, 4);
        p += 4 * nd;
        (void) rank;
        if (cls == 0) {
            n = le(p, 4);
            raw = f_.data() + p + 4;
            return have(p + 4, n) ? true : fail(err, "compact data out of the file");
        }
        if (is_undef(a)) return fail(err, "dataset has no data");
        n = total;
        raw = f_.data() + addr(a);
        return have(addr(a), n) ? true : fail(err, "data out of the file");
    }

    static double scalar(const Type &t, const uint8_t *p)
    {
        uint8_t b[8];
        for (uint32_t i = 0; i < t.size; ++i) b[i] = t.big ? p[t.size - 1 - i] : p[i];
        if (t.cls == 1) {
            if (t.size == 4) {
                float v;
                memcpy(&v, b, 4);
                return v;
            }
            double v;
            memcpy(&v, b, 8);
            return v;
        }
        uint64_t u = 0;
        for (int i = (int) t.size - 1; i >= 0; --i) u = (u << 8) | b[i];
        if (t.sgn && t.size < 8 && (u >> (8 * t.size - 1)) & 1) u |= ~0ull << (8 * t.size);
        return t.sgn ? (double) (int64_t) u : (double) u;
    }
};

static bool read_named(const std::string &path, const char *name, Dataset &ds, std::string &err)
{
    Reader r(path);
    if (!r.open(err)) return false;
    std::map<std::string, uint64_t> sets;
    if (!r.datasets(sets, err)) return false;
    std::string key = name[0] == '/' ? std::string(name + 1) : std::string(name);
    auto it = sets.find(key);
    if (it == sets.end()) {
        err = path + " has no dataset " + key;
        return false;
    }
    return r.dataset(it->second, ds, err);
}

}   // namespace h5

// The reference's solution file (channel/channel.c:889-1028): C and IC
// ([Nz][2][dimR][Nx/2]) as data_alpha/data_beta/data_ialpha/data_ibeta
// [dimR][Nz][Nx/2], the grid and the parameters as one-element datasets
// (data_Ny = 2 qpts / 3, as the reference writes it).
int h5_write_solution(const std::string &path, const Dims &d, double mpg, const cplx *C,
                      const cplx *IC, std::string &err)
{
    const size_t per = (size_t) d.dimR * d.Nz * d.Mx;
    std::vector<cplx> planes[4];
    const cplx *src[2] = {C, IC};
    for (int a = 0; a < 2; ++a)
        for (int v = 0; v < 2; ++v) {
            std::vector<cplx> &P = planes[2 * a + v];
            P.resize(per);
            for (int y = 0; y < d.dimR; ++y)
                for (int z = 0; z < d.Nz; ++z)
                    memcpy(&P[((size_t) y * d.Nz + z) * d.Mx],
                           src[a] + (((size_t) z * 2 + v) * d.dimR + y) * d.Mx, sizeof(cplx) * d.Mx);
        }
    const int32_t nx = d.Nx, ny = d.qpts * 2 / 3, nz = d.Nz;
    const double dt = d.dt, re = d.re;
    const std::vector<uint64_t> shape = {(uint64_t) d.dimR, (uint64_t) d.Nz, (uint64_t) d.Mx}, one = {1};
    const size_t pb = per * sizeof(cplx);
    std::vector<h5::DatasetSpec> sets = {
        {"data_alpha", 2, shape, planes[0].data(), pb}, {"data_beta", 2, shape, planes[1].data(), pb},
        {"data_ialpha", 2, shape, planes[2].data(), pb}, {"data_ibeta", 2, shape, planes[3].data(), pb},
        {"data_Nx", 0, one, &nx, 4}, {"data_Ny", 0, one, &ny, 4}, {"data_Nz", 0, one, &nz, 4},
        {"data_dt", 1, one, &dt, 8}, {"data_Re", 1, one, &re, 8}, {"data_mpg", 1, one, &mpg, 8}};
    return h5::write_file(path, sets, err) ? SC_OK : SC_EIO;
}

// C and IC from a solution file of the grid d (channel/channel.c:791-887)
int h5_read_solution(const std::string &path, const Dims &d, cplx *C, cplx *IC, std::string &err)
{
    const char *names[4] = {"data_alpha", "data_beta", "data_ialpha", "data_ibeta"};
    cplx *dst[2] = {C, IC};
    for (int k = 0; k < 4; ++k) {
        h5::Dataset ds;
        if (!h5::read_named(path, names[k], ds, err)) return SC_EIO;
        if (ds.kind != 2 || ds.dims.size() != 3 || ds.dims[0] != (uint64_t) d.dimR ||
            ds.dims[1] != (uint64_t) d.Nz || ds.dims[2] != (uint64_t) d.Mx) {
            err = std::string(names[k]) + " is not a [dimR][Nz][Nx/2] complex dataset of this grid";
            return SC_EIO;
        }
        const int a = k / 2, v = k % 2;
        for (int y = 0; y < d.dimR; ++y)
            for (int z = 0; z < d.Nz; ++z)
                for (int x = 0; x < d.Mx; ++x) {
                    const size_t s = 2 * (((size_t) y * d.Nz + z) * d.Mx + x);
                    dst[a][(((size_t) z * 2 + v) * d.dimR + y) * d.Mx + x] =
                        make_double2(ds.values[s], ds.values[s + 1]);
                }
    }
    return SC_OK;
}

}   // namespace sc

using namespace sc;

extern "C" {

int sc_h5_dataset_shape(const char *path, const char *dataset, long long *dims, int max_rank,
                        int *kind)
{
    if (!path || !dataset || !dims || !kind || max_rank < 0) return SC_EINVAL;
    h5::Dataset ds;
    std::string err;
    if (!h5::read_named(path, dataset, ds, err)) {
        last_error() = err;
        return SC_EIO;
    }
    if ((int) ds.dims.size() > max_rank) {
        last_error() = "dataset rank above max_rank";
        return SC_EINVAL;
    }
    for (size_t i = 0; i < ds.dims.size(); ++i) dims[i] = (long long) ds.dims[i];
    *kind = ds.kind;
    return (int) ds.dims.size();
}

int sc_h5_write_solution(const char *path, int Nx, int Ny, int Nz, double dt, double Re,
                         const sc_complex *C, const sc_complex *IC)
{
    if (!path || !C || !IC || Nx < 2 || Nx % 2 || Ny < 3 || Nz < 1 || !(Re > 0)) return SC_EINVAL;
    Dims d{};
    d.Nx = Nx;
    d.Ny = Ny;
    d.Nz = Nz;
    d.Mx = Nx / 2;
    d.dimR = Ny - 2;
    d.qpts = 3 * Ny / 2;
    d.dt = dt;
    d.re = 1.0 / Re;            // the reference stores its global re = 1/Re
    std::string err;
    const int rc = h5_write_solution(path, d, 0.0, (const cplx *) C, (const cplx *) IC, err);
    if (rc != SC_OK) last_error() = err;
    return rc;
}

long long sc_h5_read_doubles(const char *path, const char *dataset, double *out, long long cap)
{
    if (!path || !dataset || (!out && cap > 0) || cap < 0) return SC_EINVAL;
    h5::Dataset ds;
    std::string err;
    if (!h5::read_named(path, dataset, ds, err)) {
        last_error() = err;
        return SC_EIO;
    }
    if ((long long) ds.values.size() > cap) {
        last_error() = "output holds " + std::to_string(cap) + " doubles, the dataset " +
                       std::to_string(ds.values.size());
        return SC_EINVAL;
    }
    std::copy(ds.values.begin(), ds.values.end(), out);
    return (long long) ds.values.size();
}

}   // extern "C"
