// world_state_io.hpp -- WorldState files in the three formats of the
// reference's dpd_state_io.hpp:109-588, so the engine's tools read and write the
// reference's state files: text "v0", and "v1binary" / "v2binary", which keep
// the text header and store the beads as packed binary records. A file written
// here and read back by the reference (or the other way round) gives the same
// WorldState; for the same state the bytes are those the reference writes
// (text numbers in C++'s default 6-significant-digit format).
//
//   WorldState <v0|v1binary|v2binary> <bead types> <polymer types> <beads> <polymers>
//   T <t> <dt>
//   Lambda <lambda>
//   Origin <x> <y> <z>
//   Box <x> <y> <z>
//   Seed <seed>
//   ConservativeStrength <i> <a_i0> ... ; DissipativeStrength <i> <g_i0> ...
//   BeadType <id> <name> <r> [stationary]
//   PolymerType <id> <name> <n bead types> <n bonds> <n bond pairs>
//     BeadTypeIndices <t0> ... ; Bond <head> <tail> <kappa> <r0> ;
//     BondPair <head bond> <tail bond> <kappa> <theta0>
//   v0:     B <bead id> <polymer id> <polymer type> <offset>  <x> <y> <z>  <v...>  <f...>
//   binary: a line "BinaryData", a NUL byte, then one record per polymer.
//
// Lines whose first word starts with '#' and empty lines are skipped. Bead
// positions are kept relative to the origin; text files hold them with the
// origin added, binary files without. Parameters are numbers or expressions
// (pdpd/expression.hpp; no spaces inside, as fields are split on spaces).
// ".gz" and ".bz2" files go through gzip / bzip2.
//
// Binary polymer record (little-endian):
//   u16  bits 0-12 polymer type; bit 13: polymer id is not previous + 1;
//        bit 14: first bead id is not the previous bead id + 1;
//        bit 15: bead ids inside the polymer are not consecutive
//   [u32 polymer id - previous polymer id]            if bit 13
//   per bead, in polymer-offset order:
//     [u32 bead id - previous bead id]                if bit 14 (first bead) or bit 15 (others)
//     u32 x3  position relative to the origin in units of 2^-16
//     v, f    v2: 4 bytes each (pack_half3); v1: 8 bytes each (int16 exponent e,
//             then int16 m x3; component = m * 2^(e-15))
// "Previous" ids start at 0xFFFFFFFF, so sequential numbering from 0 needs no ids.
#pragma once

#include "pdpd/world_state.hpp"

#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <istream>
#include <memory>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace pdpd_io {

// A line split on spaces, with the reference's typed accessors.
struct Fields {
    std::vector<std::string> parts;
    int line_no = 0;

    const std::string &at(size_t i) const
    {
        if (i >= parts.size()) throw std::runtime_error("missing field " + std::to_string(i) + " at line " + std::to_string(line_no));
        return parts[i];
    }
    // A number (strtod, so also a negative one, which the expression language
    // has no literal for), else an expression.
    Parameter param(size_t i) const
    {
        const std::string &t = at(i);
        char *e = nullptr;
        const double v = std::strtod(t.c_str(), &e);
        if (!t.empty() && e && *e == 0) return Parameter(v);
        try {
            return Parameter(t);
        } catch (const std::runtime_error &e) {
            throw std::runtime_error(std::string(e.what()) + " (parameter '" + at(i) + "' at line " +
                                     std::to_string(line_no) + ")");
        }
    }
    double real(size_t i) const
    {
        const std::string &s = at(i);
        char *end = nullptr;
        const double v = std::strtod(s.c_str(), &end);
        if (end != s.c_str() + s.size())
            throw std::runtime_error("not a number (parameter expressions are not supported): '" + s + "' at line " +
                                     std::to_string(line_no));
        return v;
    }
    uint64_t uint(size_t i) const
    {
        const std::string &s = at(i);
        char *end = nullptr;
        const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
        if (s.empty() || end != s.c_str() + s.size() || s[0] == '-')
            throw std::runtime_error("not an unsigned integer: '" + s + "' at line " + std::to_string(line_no));
        return v;
    }
};

// Next line with content whose first word is `prefix` (n = expected word count, -1 any).
inline Fields expect(std::istream &src, const char *prefix, int n, int &line_no)
{
    std::string line;
    while (std::getline(src, line)) {
        line_no++;
        Fields f;
        f.line_no = line_no;
        std::string word;
        for (char c : line) {
            if (c == ' ' || c == '\r' || c == '\t') {
                if (!word.empty()) { f.parts.push_back(word); word.clear(); }
            } else {
                word += c;
            }
        }
        if (!word.empty()) f.parts.push_back(word);
        if (f.parts.empty() || f.parts[0][0] == '#') continue;
        if (f.parts[0] != prefix)
            throw std::runtime_error("expected '" + std::string(prefix) + "' but got '" + f.parts[0] + "' at line " +
                                     std::to_string(line_no));
        if (n >= 0 && (int)f.parts.size() != n)
            throw std::runtime_error("expected " + std::to_string(n) + " fields for '" + prefix + "' at line " +
                                     std::to_string(line_no));
        return f;
    }
    throw std::runtime_error("expected '" + std::string(prefix) + "' but reached the end of the file");
}

// The reference prints a parameter as its expression text, a constant with
// std::to_string (expression.hpp:431-435).
inline std::string param(const Parameter &p) { return p.as_string(); }

// ---- binary records -------------------------------------------------------

// A velocity or force in 4 bytes (the v2binary form; CVec3Half of the
// reference's cvec3.hpp): bits 0-4 hold e+15 for a common scale 2^e with e in
// [-14, 15] (0 = the zero vector), bits 5-13 / 14-22 / 23-31 the components as
// 9-bit two's-complement m, component = m * 2^(e-8). e is the largest binary
// exponent of the components (frexp), m = round(c * 2^(8-e)) with ties away from
// zero; if a component rounds to +-256, e goes up by one. Subnormal components
// count as zero; vectors below 2^-15 are stored as zero; NaN, infinity and
// magnitudes of 2^15 or more are errors.
inline uint32_t pack_half3(const double p[3])
{
    double c[3];
    bool zero = true;
    for (int d = 0; d < 3; d++) {
        const int cls = std::fpclassify(p[d]);
        if (cls == FP_NAN || cls == FP_INFINITE) throw std::runtime_error("cannot pack NaN or infinity into a v2binary vector");
        c[d] = cls == FP_NORMAL ? p[d] : 0.0;
        zero = zero && c[d] == 0.0;
    }
    if (zero) return 0;
    int e = INT_MIN;
    for (int d = 0; d < 3; d++)
        if (c[d] != 0.0) {
            int k;
            std::frexp(c[d], &k);
            e = std::max(e, k);
        }
    if (e > 15) throw std::runtime_error("vector too large for a v2binary record (magnitude 65536 or more)");
    if (e < -14) return 0;
    double m[3];
    bool carry = false;
    for (int d = 0; d < 3; d++) {
        m[d] = std::round(std::ldexp(c[d], 8 - e));
        carry = carry || std::fabs(m[d]) == 256.0;
    }
    if (carry) {
        if (++e > 15) throw std::runtime_error("vector too large for a v2binary record (magnitude 65536 or more)");
        for (int d = 0; d < 3; d++) m[d] = std::round(std::ldexp(c[d], 8 - e));
    }
    uint32_t bits = uint32_t(e + 15);
    for (int d = 0; d < 3; d++) bits |= (uint32_t(int32_t(m[d])) & 0x1ffu) << (5 + 9 * d);
    return bits;
}

inline vec3r_t unpack_half3(uint32_t bits)
{
    const int e = int(bits & 31u);
    vec3r_t r;
    for (int d = 0; d < 3; d++) {
        int32_t m = int32_t((bits >> (5 + 9 * d)) & 0x1ffu);
        if (m & 0x100) m -= 0x200;
        r[d] = std::ldexp((double)m, e - 15 - 8);
    }
    return r;
}

// A velocity or force in 8 bytes (the v1binary form; CVec3): int16 e, then the
// components as int16 m, component = m * 2^(e-15). Read only: the reference
// writes v2binary.
inline vec3r_t unpack_c3(const int16_t w[4])
{
    vec3r_t r;
    for (int d = 0; d < 3; d++) r[d] = std::ldexp((double)w[1 + d], w[0] - 15);
    return r;
}

// A position coordinate relative to the origin as u32 in units of 2^-16,
// rounded to nearest (ties away from zero); a coordinate that rounds up to the
// box edge wraps to 0.
inline uint32_t pack_pos(double x, double box)
{
    double r = std::round(x * 65536.0);
    if (r == box * 65536.0) r = 0;
    if (!(r >= 0 && r < box * 65536.0)) throw std::runtime_error("Bead is out of bounds.");
    return (uint32_t)r;
}

inline void put_u16(std::ostream &o, uint16_t v)
{
    const char b[2] = {char(v & 0xff), char(v >> 8)};
    o.write(b, 2);
}
inline void put_u32(std::ostream &o, uint32_t v)
{
    const char b[4] = {char(v & 0xff), char((v >> 8) & 0xff), char((v >> 16) & 0xff), char(v >> 24)};
    o.write(b, 4);
}
inline uint32_t get_le(std::istream &i, int n)
{
    unsigned char b[4] = {0, 0, 0, 0};
    i.read((char *)b, n);
    if (i.gcount() != n) throw std::runtime_error("binary bead data ends early");
    return uint32_t(b[0]) | uint32_t(b[1]) << 8 | uint32_t(b[2]) << 16 | uint32_t(b[3]) << 24;
}

inline void write_binary_beads(std::ostream &o, const WorldState &s)
{
    uint32_t prev_polymer = 0xFFFFFFFFu, prev_bead = 0xFFFFFFFFu;
    for (const Polymer &p : s.polymers) {
        if (p.bead_ids.empty()) throw std::runtime_error("polymer " + std::to_string(p.polymer_id) + " has no beads");
        if (p.polymer_type >= (1u << 13)) throw std::runtime_error("polymer type does not fit the 13-bit binary field");
        const bool jump_polymer = p.polymer_id != prev_polymer + 1;
        const bool jump_first = p.bead_ids[0] != prev_bead + 1;
        bool jump_inside = false;
        for (size_t k = 1; k < p.bead_ids.size(); k++) jump_inside = jump_inside || p.bead_ids[k] != p.bead_ids[k - 1] + 1;
        put_u16(o, uint16_t(p.polymer_type | uint32_t(jump_polymer) << 13 | uint32_t(jump_first) << 14 |
                            uint32_t(jump_inside) << 15));
        if (jump_polymer) put_u32(o, p.polymer_id - prev_polymer);
        prev_polymer = p.polymer_id;
        for (size_t k = 0; k < p.bead_ids.size(); k++) {
            const uint32_t id = p.bead_ids[k];
            if (k == 0 ? jump_first : jump_inside) put_u32(o, id - prev_bead);
            prev_bead = id;
            const Bead &b = s.beads.at(id);
            for (int d = 0; d < 3; d++) put_u32(o, pack_pos(b.x[d], s.box[d]));
            put_u32(o, pack_half3(&b.v.x[0]));
            put_u32(o, pack_half3(&b.f.x[0]));
        }
    }
}

inline void read_binary_beads(std::istream &in, WorldState &s, bool v2)
{
    uint32_t prev_polymer = 0xFFFFFFFFu, prev_bead = 0xFFFFFFFFu;
    for (size_t i = 0; i < s.polymers.size(); i++) {
        const uint32_t h = get_le(in, 2);
        const uint32_t type = h & 0x1fffu;
        const bool jump_first = (h >> 14) & 1u, jump_inside = (h >> 15) & 1u;
        const uint32_t pid = prev_polymer + (((h >> 13) & 1u) ? get_le(in, 4) : 1u);
        prev_polymer = pid;
        if (pid >= s.polymers.size()) throw std::runtime_error("Polymer id out of range.");
        if (type >= s.polymer_types.size()) throw std::runtime_error("polymer type out of range in binary data");
        Polymer &poly = s.polymers[pid];
        if (poly.polymer_id != uint32_t(-1)) throw std::runtime_error("Polymer appears twice.");
        poly.polymer_id = pid;
        poly.polymer_type = type;
        const PolymerType &pt = s.polymer_types[type];
        for (size_t k = 0; k < pt.bead_types.size(); k++) {
            const uint32_t id = prev_bead + ((k == 0 ? jump_first : jump_inside) ? get_le(in, 4) : 1u);
            prev_bead = id;
            if (id >= s.beads.size()) throw std::runtime_error("bead id out of range in binary data");
            Bead &b = s.beads[id];
            if (b.bead_id != uint32_t(-1)) throw std::runtime_error("Bead appears twice.");
            poly.bead_ids.push_back(id);
            b.bead_id = id;
            b.bead_type = (uint8_t)pt.bead_types[k];
            b.is_monomer = pt.bead_types.size() == 1;
            b.polymer_id = pid;
            b.polymer_offset = (uint8_t)k;
            b.polymer_type = (uint8_t)type;
            for (int d = 0; d < 3; d++) {
                b.x[d] = std::ldexp((double)get_le(in, 4), -16);
                if (!(b.x[d] < s.box[d])) throw std::runtime_error("Bead was out of box.");
            }
            if (v2) {
                b.v = unpack_half3(get_le(in, 4));
                b.f = unpack_half3(get_le(in, 4));
            } else {
                for (vec3r_t *dst : {&b.v, &b.f}) {
                    int16_t w[4];
                    for (int q = 0; q < 4; q++) w[q] = (int16_t)(uint16_t)get_le(in, 2);
                    *dst = unpack_c3(w);
                }
            }
        }
    }
}

// binary = true writes v2binary (as the reference's write_world_state(..., true));
// precise = true writes the text numbers with 17 significant digits instead of 6.
inline void write_world_state(std::ostream &dst, const WorldState &s, bool binary = false, bool precise = false)
{
    std::ostringstream o;            // default stream format: 6 significant digits, like the reference
    if (precise) o << std::setprecision(17);
    const size_t nt = s.bead_types.size();
    o << (binary ? "WorldState v2binary " : "WorldState v0 ") << nt << " " << s.polymer_types.size() << " " << s.beads.size() << " " << s.polymers.size()
      << "\n";
    o << "T " << s.t << " " << s.dt << "\n";
    o << "Lambda " << s.lambda << "\n";
    o << "Origin " << s.origin[0] << " " << s.origin[1] << " " << s.origin[2] << "\n";
    o << "Box " << s.box[0] << " " << s.box[1] << " " << s.box[2] << "\n";
    o << "Seed " << s.seed << "\n";
    for (size_t i = 0; i < nt; i++) {
        o << "ConservativeStrength " << i;
        for (size_t j = 0; j < nt; j++) o << " " << param(s.interactions[i * nt + j].conservative);
        o << "\n";
    }
    for (size_t i = 0; i < nt; i++) {
        o << "DissipativeStrength " << i;
        for (size_t j = 0; j < nt; j++) o << " " << param(s.interactions[i * nt + j].dissipative);
        o << "\n";
    }
    o << "\n# BeadTypes\n";
    for (const BeadType &b : s.bead_types) {
        o << "BeadType " << b.id << " " << b.name << " " << b.r;
        if (b.stationary) o << " stationary";
        o << "\n";
    }
    o << "\n# polymerTypes\n";
    for (const PolymerType &p : s.polymer_types) {
        o << "PolymerType " << p.polymer_id << " " << p.name << " " << p.bead_types.size() << " " << p.bonds.size()
          << " " << p.bond_pairs.size() << "\n";
        o << "BeadTypeIndices";
        for (unsigned t : p.bead_types) o << " " << t;
        o << "\n";
        for (const Bond &b : p.bonds)
            o << "Bond " << b.bead_offset_head << " " << b.bead_offset_tail << " " << param(b.kappa) << " "
              << param(b.r0) << "\n";
        for (const BondPair &b : p.bond_pairs)
            o << "BondPair " << b.bond_offset_head << " " << b.bond_offset_tail << " " << param(b.kappa) << " "
              << param(b.theta0) << "\n";
        o << "\n";
    }
    if (binary) {
        o << "\nBinaryData\n";
        o.put('\0');
        write_binary_beads(o, s);
    } else {
        o << "\n# Beads\n";
        for (const Bead &b : s.beads) {
            o << "B " << b.bead_id << " " << b.polymer_id << " " << s.polymers.at(b.polymer_id).polymer_type << " "
              << (unsigned)b.polymer_offset << " ";
            for (int d = 0; d < 3; d++) o << " " << b.x[d] + s.origin[d];
            o << " ";
            for (int d = 0; d < 3; d++) o << " " << b.v[d];
            o << " ";
            for (int d = 0; d < 3; d++) o << " " << b.f[d];
            o << "\n";
        }
        o << "\n";
    }
    dst << o.str();
    if (!dst) throw std::runtime_error("error writing the world state");
}

inline WorldState read_world_state(std::istream &src)
{
    int ln = 0;
    WorldState s;
    Fields h = expect(src, "WorldState", 6, ln);
    const bool binary = h.at(1) == "v1binary" || h.at(1) == "v2binary";
    if (!binary && h.at(1) != "v0")
        throw std::runtime_error("WorldState format '" + h.at(1) + "' is not supported (v0, v1binary, v2binary)");
    const size_t nt = h.uint(2), npt = h.uint(3), nb = h.uint(4), np = h.uint(5);
    Fields f = expect(src, "T", 3, ln);
    s.t = (long)f.real(1);
    s.dt = f.real(2);
    s.lambda = expect(src, "Lambda", 2, ln).real(1);
    f = expect(src, "Origin", 4, ln);
    for (int d = 0; d < 3; d++) s.origin[d] = f.real(d + 1);
    f = expect(src, "Box", 4, ln);
    for (int d = 0; d < 3; d++) s.box[d] = f.real(d + 1);
    s.seed = expect(src, "Seed", 2, ln).uint(1);
    s.interactions.assign(nt * nt, InteractionStrength{});
    for (size_t i = 0; i < nt; i++) {
        f = expect(src, "ConservativeStrength", (int)nt + 2, ln);
        for (size_t j = 0; j < nt; j++) s.interactions[i * nt + j].conservative = f.param(j + 2);
    }
    for (size_t i = 0; i < nt; i++) {
        f = expect(src, "DissipativeStrength", (int)nt + 2, ln);
        for (size_t j = 0; j < nt; j++) s.interactions[i * nt + j].dissipative = f.param(j + 2);
    }
    s.bead_types.resize(nt);
    for (size_t i = 0; i < nt; i++) {
        f = expect(src, "BeadType", -1, ln);
        if (f.parts.size() != 4 && f.parts.size() != 5)
            throw std::runtime_error("expected 4 or 5 fields for BeadType at line " + std::to_string(ln));
        BeadType &b = s.bead_types.at(i);
        b.id = (uint32_t)f.uint(1);
        b.name = f.at(2);
        b.r = f.real(3);
        b.stationary = false;
        if (f.parts.size() == 5) {
            if (f.at(4) != "stationary") throw std::runtime_error("unknown BeadType flag '" + f.at(4) + "'");
            b.stationary = true;
        }
    }
    s.polymer_types.resize(npt);
    std::vector<bool> seen_pt(npt, false);
    for (size_t i = 0; i < npt; i++) {
        f = expect(src, "PolymerType", 6, ln);
        const size_t id = f.uint(1);
        if (id >= npt || seen_pt[id]) throw std::runtime_error("bad or repeated polymer type id at line " + std::to_string(ln));
        seen_pt[id] = true;
        PolymerType &p = s.polymer_types[id];
        p.polymer_id = (uint32_t)id;
        p.name = f.at(2);
        p.bead_types.resize(f.uint(3));
        p.bonds.resize(f.uint(4));
        p.bond_pairs.resize(f.uint(5));
        Fields t = expect(src, "BeadTypeIndices", (int)p.bead_types.size() + 1, ln);
        for (size_t k = 0; k < p.bead_types.size(); k++) {
            p.bead_types[k] = (unsigned)t.uint(k + 1);
            if (p.bead_types[k] >= nt) throw std::runtime_error("bead type index out of range at line " + std::to_string(ln));
        }
        for (Bond &b : p.bonds) {
            Fields g = expect(src, "Bond", 5, ln);
            b.bead_offset_head = (uint32_t)g.uint(1);
            b.bead_offset_tail = (uint32_t)g.uint(2);
            b.kappa = g.param(3);
            b.r0 = g.param(4);
            if (b.bead_offset_head >= p.bead_types.size() || b.bead_offset_tail >= p.bead_types.size())
                throw std::runtime_error("bond bead offset out of range at line " + std::to_string(ln));
        }
        for (BondPair &b : p.bond_pairs) {
            Fields g = expect(src, "BondPair", 5, ln);
            b.bond_offset_head = (uint32_t)g.uint(1);
            b.bond_offset_tail = (uint32_t)g.uint(2);
            b.kappa = g.param(3);
            b.theta0 = g.param(4);
            if (b.bond_offset_head >= p.bonds.size() || b.bond_offset_tail >= p.bonds.size())
                throw std::runtime_error("bond pair offset out of range at line " + std::to_string(ln));
        }
    }
    s.polymers.assign(np, Polymer{});
    s.beads.assign(nb, Bead{});
    if (binary) {
        expect(src, "BinaryData", 1, ln);
        for (;;) {
            const int c = src.get();
            if (c == EOF) throw std::runtime_error("Missing null before binary data.");
            if (c == 0) break;
            if (!std::isspace(c)) throw std::runtime_error("Unexpected char before null char.");
        }
        read_binary_beads(src, s, h.at(1) == "v2binary");
        for (const Bead &b : s.beads)
            if (b.bead_id == uint32_t(-1)) throw std::runtime_error("a bead is missing from the binary data");
        return s;
    }
    for (size_t i = 0; i < nb; i++) {
        f = expect(src, "B", 14, ln);
        Bead b;
        b.bead_id = (uint32_t)f.uint(1);
        b.polymer_id = (uint32_t)f.uint(2);
        b.polymer_type = (uint8_t)f.uint(3);
        b.polymer_offset = (uint8_t)f.uint(4);
        if (b.bead_id >= nb || b.polymer_id >= np || f.uint(3) >= npt)
            throw std::runtime_error("bead, polymer or polymer type id out of range at line " + std::to_string(ln));
        const PolymerType &pt = s.polymer_types[f.uint(3)];
        if (f.uint(4) >= pt.bead_types.size()) throw std::runtime_error("polymer offset out of range at line " + std::to_string(ln));
        Polymer &poly = s.polymers[b.polymer_id];
        if (poly.polymer_id == uint32_t(-1)) {
            poly.polymer_id = b.polymer_id;
            poly.polymer_type = (uint32_t)f.uint(3);
            poly.bead_ids.assign(pt.bead_types.size(), uint32_t(-1));
        } else if (poly.polymer_type != f.uint(3)) {
            throw std::runtime_error("polymer type changes within a polymer at line " + std::to_string(ln));
        }
        if (poly.bead_ids[b.polymer_offset] != uint32_t(-1))
            throw std::runtime_error("bead repeated within a polymer at line " + std::to_string(ln));
        poly.bead_ids[b.polymer_offset] = b.bead_id;
        b.is_monomer = pt.bead_types.size() == 1;
        b.bead_type = (uint8_t)pt.bead_types[b.polymer_offset];
        for (int d = 0; d < 3; d++) {
            b.x[d] = f.real(d + 5) - s.origin[d];
            if (b.x[d] == s.box[d]) b.x[d] = 0;            // a coordinate rounded up to the box edge on output
            b.v[d] = f.real(d + 8);
            b.f[d] = f.real(d + 11);
        }
        if (s.beads[b.bead_id].bead_id != uint32_t(-1))
            throw std::runtime_error("bead id repeated at line " + std::to_string(ln));
        s.beads[b.bead_id] = b;
    }
    for (const Polymer &p : s.polymers)
        for (uint32_t id : p.bead_ids)
            if (id == uint32_t(-1)) throw std::runtime_error("a polymer is missing beads");
    return s;
}

// Files: ".gz" names go through gzip, ".bz2" names through bzip2; anything
// else is a plain file.
inline std::string compressor(const std::string &name)
{
    auto ends = [&](const char *x) {
        const size_t n = strlen(x);
        return name.size() > n && name.compare(name.size() - n, n, x) == 0;
    };
    return ends(".gz") ? "gzip" : ends(".bz2") ? "bzip2" : "";
}

inline std::string shell_quote(const std::string &s)
{
    std::string q = "'";
    for (char c : s) q += (c == '\'') ? std::string("'\\''") : std::string(1, c);
    return q + "'";
}

inline WorldState read_world_state(const std::string &name)
{
    const std::string z = compressor(name);
    if (!z.empty()) {
        FILE *f = popen((z + " -dc " + shell_quote(name)).c_str(), "r");
        if (!f) throw std::runtime_error("could not run " + z + " for " + name);
        std::string data;
        char buf[1 << 16];
        size_t n;
        while ((n = fread(buf, 1, sizeof(buf), f)) > 0) data.append(buf, n);
        if (pclose(f) != 0) throw std::runtime_error(z + " could not read " + name);
        std::istringstream in(data);
        return read_world_state(in);
    }
    std::ifstream in(name, std::ios::binary);
    if (!in) throw std::runtime_error("could not open " + name);
    return read_world_state(in);
}

inline void write_world_state(const std::string &name, const WorldState &s, bool binary = false, bool precise = false)
{
    std::ostringstream data;
    write_world_state(data, s, binary, precise);
    const std::string t = data.str();
    const std::string z = compressor(name);
    if (!z.empty()) {
        FILE *f = popen((z + " -c > " + shell_quote(name)).c_str(), "w");
        if (!f) throw std::runtime_error("could not run " + z + " for " + name);
        const bool ok = fwrite(t.data(), 1, t.size(), f) == t.size();
        if (pclose(f) != 0 || !ok) throw std::runtime_error("error writing " + name);
        return;
    }
    std::ofstream out(name, std::ios::binary);
    out.write(t.data(), (std::streamsize)t.size());
    if (!out) throw std::runtime_error("could not write " + name);
}

}  // namespace pdpd_io
