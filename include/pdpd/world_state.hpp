// world_state.hpp -- the WorldState data model of poets-dpd, restated for this
// repository so the engines here compile without the reference tree.
//
// Names, fields and meanings follow include/dpd/core/dpd_state.hpp:22-264 and
// include/dpd/core/vec3.hpp of poets-dpd, because code written against that
// interface (tools, tests, the HIP engine) must compile against either copy.
// Parameters are constants or expressions (include/dpd/core/expression.hpp;
// pdpd/expression.hpp here); the HIP engine supports DPD strengths that vary
// with the cell position (variables x, y, z, ux, uy, uz) and rejects others.
#pragma once

#include "pdpd/expression.hpp"

#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <ostream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

template <class T>
struct vec3g_t {
    std::array<T, 3> x{};

    vec3g_t() : x{0, 0, 0} {}
    vec3g_t(T a, T b, T c) : x{a, b, c} {}
    template <class U>
    vec3g_t(const vec3g_t<U> &o) : x{(T)o[0], (T)o[1], (T)o[2]} {}

    T operator[](size_t i) const { return x[i]; }
    T &operator[](size_t i) { return x[i]; }

    template <class U>
    void assign(const U src[3]) { for (int i = 0; i < 3; i++) x[i] = (T)src[i]; }
    template <class U>
    void extract(U dst[3]) const { for (int i = 0; i < 3; i++) dst[i] = (U)x[i]; }
    void clear() { x = {0, 0, 0}; }

    bool operator==(const vec3g_t &o) const { return x == o.x; }
    bool operator!=(const vec3g_t &o) const { return x != o.x; }
    vec3g_t operator+(const vec3g_t &o) const { return {x[0] + o[0], x[1] + o[1], x[2] + o[2]}; }
    vec3g_t operator-(const vec3g_t &o) const { return {x[0] - o[0], x[1] - o[1], x[2] - o[2]}; }
    vec3g_t operator*(T s) const { return {x[0] * s, x[1] * s, x[2] * s}; }
    vec3g_t operator/(T s) const { return {x[0] / s, x[1] / s, x[2] / s}; }
    vec3g_t operator-() const { return {-x[0], -x[1], -x[2]}; }
    vec3g_t &operator+=(const vec3g_t &o) { *this = *this + o; return *this; }
    vec3g_t &operator-=(const vec3g_t &o) { *this = *this - o; return *this; }
    T dot(const vec3g_t &o) const { return x[0] * o[0] + x[1] * o[1] + x[2] * o[2]; }
    T l2_norm() const { return std::sqrt(dot(*this)); }
};

template <class T>
inline std::ostream &operator<<(std::ostream &o, const vec3g_t<T> &v)
{
    return o << "(" << v[0] << "," << v[1] << "," << v[2] << ")";
}

using vec3r_t = vec3g_t<double>;
using vec3f_t = vec3g_t<float>;
using vec3i_t = vec3g_t<int32_t>;

// A constant parameter (the constant case of expression.hpp's Parameter).
// A constant, or an expression of variables (pdpd/expression.hpp; the
// reference's Parameter, expression.hpp:328-431). Using an expression as a
// constant throws; engines evaluate spatial ones per cell.
struct Parameter {
    double m_value = 0.0;
    pdpd_expr::NodePtr m_expr;
    Parameter() = default;
    Parameter(double v) : m_value(v) {}
    explicit Parameter(const std::string &src) { set(pdpd_expr::parse(src)); }
    Parameter &operator=(double v) { m_value = v; m_expr.reset(); return *this; }
    operator double() const
    {
        if (m_expr)
            throw std::runtime_error("Attempt to use expression with variables as a constant. Expr=." + m_expr->text());
        return m_value;
    }
    bool operator==(const Parameter &o) const
    {
        return m_expr ? (o.m_expr && m_expr->text() == o.m_expr->text()) : (!o.m_expr && m_value == o.m_value);
    }
    bool is_constant() const { return !m_expr; }
    void collect_variables(std::unordered_set<std::string> &out) const { if (m_expr) m_expr->variables(out); }
    double evaluate(const std::unordered_map<std::string, double> &b) const { return m_expr ? m_expr->eval(b) : m_value; }
    std::string as_string() const { return m_expr ? m_expr->text() : std::to_string(m_value); }

private:
    void set(pdpd_expr::NodePtr e)
    {
        if (e->op == 'c') m_value = e->value;
        else m_expr = std::move(e), m_value = 0.0;
    }
};

// 32-bit bead identity (dpd_state.hpp:18-35):
//   bbbb 1ppp pppp pppp  pppp pppp pppp pppp   monomer, 27-bit id
//   bbbb 0ooo oooo pppp  pppp pppp pppp pppp   polymer bead, 20-bit id, 7-bit offset
struct BeadHash {
    static const unsigned MONOMER_ID_BITS = 27;
    static const unsigned POLYMER_ID_BITS = 20;
    static const unsigned POLYMER_ID_MASK = (1u << POLYMER_ID_BITS) - 1;
    static const unsigned BEAD_TYPE_BITS = 4;
    static const unsigned BEAD_TYPE_SHIFT = 28;
    static const unsigned MONOMER_FLAG_SHIFT = 27;
    static const unsigned POLYMER_OFFSET_BITS = 7;
    static const unsigned POLYMER_OFFSET_MASK = (1u << POLYMER_OFFSET_BITS) - 1;
    static const unsigned POLYMER_OFFSET_SHIFT = 20;

    uint32_t hash = 0;
    BeadHash() = default;
    explicit BeadHash(uint32_t h) : hash(h) {}

    static BeadHash construct(uint32_t bead_type, bool is_monomer, uint32_t polymer_id, uint32_t polymer_offset)
    {
        uint32_t h = (polymer_offset << POLYMER_OFFSET_SHIFT) | polymer_id;
        h |= uint32_t(is_monomer) << MONOMER_FLAG_SHIFT;
        h |= bead_type << BEAD_TYPE_SHIFT;
        return BeadHash(h);
    }
    bool is_monomer() const { return (hash >> MONOMER_FLAG_SHIFT) & 1u; }
    uint32_t get_bead_type() const { return hash >> BEAD_TYPE_SHIFT; }
    uint32_t get_polymer_offset() const
    {
        return is_monomer() ? 0 : ((hash >> POLYMER_OFFSET_SHIFT) & POLYMER_OFFSET_MASK);
    }
    uint32_t get_polymer_id() const { return hash & POLYMER_ID_MASK; }
    // The identity without the type bits (dpd_state.hpp:40,103-104).
    static const unsigned REDUCED_HASH_MASK = 0xFFFFFFFFu >> BEAD_TYPE_BITS;
    BeadHash reduced_hash() const { return BeadHash(hash & REDUCED_HASH_MASK); }
    bool operator==(const BeadHash &o) const { return hash == o.hash; }
    bool operator!=(const BeadHash &o) const { return hash != o.hash; }
};

static const float MIN_DISTANCE_CUTOFF = 0.000000001f;
static const float MIN_DISTANCE_CUTOFF_SQR = MIN_DISTANCE_CUTOFF * MIN_DISTANCE_CUTOFF;

struct Bead {
    uint32_t bead_id = uint32_t(-1);
    uint32_t polymer_id = uint32_t(-1);
    uint8_t polymer_offset = uint8_t(-1);
    uint8_t bead_type = uint8_t(-1);
    uint8_t polymer_type = uint8_t(-1);
    bool is_monomer = false;
    vec3r_t x, v, f;

    uint32_t get_bead_type() const { return bead_type; }
    BeadHash get_hash_code() const { return BeadHash::construct(bead_type, is_monomer, polymer_id, polymer_offset); }
};

struct BeadType {
    std::string name;
    double r = 0.5;
    uint32_t id = uint32_t(-1);
    bool stationary = false;
};

struct Bond {
    Parameter kappa, r0;
    uint32_t bead_offset_head = uint32_t(-1);
    uint32_t bead_offset_tail = uint32_t(-1);
};

struct BondPair {
    Parameter kappa, theta0;
    uint32_t bond_offset_head = uint32_t(-1);
    uint32_t bond_offset_tail = uint32_t(-1);
};

struct PolymerType {
    std::string name;
    std::vector<unsigned> bead_types;
    std::vector<Bond> bonds;
    std::vector<BondPair> bond_pairs;
    uint32_t polymer_id = uint32_t(-1);
};

struct Polymer {
    std::vector<uint32_t> bead_ids;
    uint32_t polymer_id = uint32_t(-1);
    uint32_t polymer_type = uint32_t(-1);
};

struct InteractionStrength {
    Parameter conservative, dissipative;
};

struct WorldState {
    vec3r_t origin = {0, 0, 0};
    vec3r_t box = {4, 4, 4};
    double lambda = 0.5;
    long t = 0;
    double dt = 0.05;
    uint64_t seed = 1;
    std::vector<InteractionStrength> interactions;   // bead_types^2, symmetric
    std::vector<BeadType> bead_types;
    std::vector<PolymerType> polymer_types;
    std::vector<Polymer> polymers;
    std::vector<Bead> beads;
};
