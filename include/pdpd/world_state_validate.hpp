// world_state_validate.hpp -- the consistency checks the reference's tools run
// on a WorldState before attaching it to an engine (dpd_state_validator.hpp:29-175,
// called by src/run_world.cpp:156-157 with the engine's GetMaxBondLength()).
// A failed check throws std::runtime_error whose message starts with
// "WorldState validation failed : " and names the check; the reference ends the
// process with exit(1) for velocity and force limits, this throws for those too.
#pragma once

#include "pdpd/world_state.hpp"

#include <cmath>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

namespace pdpd_io {

inline void validate_world_state(const WorldState &s, double max_r = 1.0)
{
    auto fail = [](const std::string &what) { throw std::runtime_error("WorldState validation failed : " + what); };
    auto need = [&](bool ok, const char *what) {
        if (!ok) fail(what);
    };

    need(0 <= s.lambda && s.lambda <= 1, "0 <= lambda <= 1");
    need(0 < s.dt && s.dt < 1, "0 < dt < 1");
    need(s.t >= 0, "t >= 0");
    for (int d = 0; d < 3; d++) {
        need(s.box[d] == std::round(s.box[d]), "box sides are integers");
        need(s.box[d] >= 4, "box sides are at least 4");
    }

    const size_t nt = s.bead_types.size();
    std::unordered_set<std::string> names;
    for (size_t i = 0; i < nt; i++) {
        const BeadType &bt = s.bead_types[i];
        need(bt.id == i, "bead type ids are 0, 1, 2, ...");
        need(names.insert(bt.name).second, "bead type names are unique");
        need(0 < bt.r && bt.r <= 0.5, "0 < bead radius <= 0.5");
    }
    need(s.interactions.size() == nt * nt, "interactions has bead_types^2 entries");
    for (size_t i = 0; i < nt; i++)
        for (size_t j = 0; j < nt; j++) {
            // constant strengths only; expressions are evaluated per cell by the engine
            const InteractionStrength &a = s.interactions[i * nt + j], &b = s.interactions[j * nt + i];
            if (a.conservative.is_constant()) {
                need(a.conservative == b.conservative, "conservative strengths are symmetric");
                need(0 <= (double)a.conservative, "conservative strengths are >= 0");
            }
            if (a.dissipative.is_constant()) {
                need(a.dissipative == b.dissipative, "dissipative strengths are symmetric");
                need(0 <= (double)a.dissipative, "dissipative strengths are >= 0");
            }
        }

    names.clear();
    for (size_t i = 0; i < s.polymer_types.size(); i++) {
        const PolymerType &pt = s.polymer_types[i];
        need(pt.polymer_id == i, "polymer type ids are 0, 1, 2, ...");
        need(names.insert(pt.name).second, "polymer type names are unique");
        for (unsigned t : pt.bead_types) need(t < nt, "polymer bead types exist");
        for (const Bond &b : pt.bonds) {
            need(b.bead_offset_head < pt.bead_types.size() && b.bead_offset_tail < pt.bead_types.size(),
                 "bond offsets lie inside the polymer");
            need(b.bead_offset_head != b.bead_offset_tail, "a bond joins two different beads");
            need(0 < (double)b.kappa, "bond kappa > 0");
            need(0 < (double)b.r0 && (double)b.r0 < 1, "0 < bond r0 < 1");
        }
        for (const BondPair &bp : pt.bond_pairs) {
            need(bp.bond_offset_head < pt.bonds.size() && bp.bond_offset_tail < pt.bonds.size(),
                 "bond pair offsets name bonds of the polymer");
            need(bp.bond_offset_head != bp.bond_offset_tail, "a bond pair joins two different bonds");
            need(pt.bonds[bp.bond_offset_head].bead_offset_tail == pt.bonds[bp.bond_offset_tail].bead_offset_head,
                 "the head bond's tail bead is the tail bond's head bead");
            need(0 < (double)bp.kappa, "bond pair kappa > 0");
            need(0 <= (double)bp.theta0 && (double)bp.theta0 <= 3.1416, "0 <= theta0 <= pi");
        }
    }

    std::vector<char> polymer_seen(s.polymers.size(), 0), bead_seen(s.beads.size(), 0);
    for (const Polymer &p : s.polymers) {
        need(p.polymer_id < s.polymers.size() && !polymer_seen[p.polymer_id], "polymer ids are unique");
        polymer_seen[p.polymer_id] = 1;
        need(p.polymer_type < s.polymer_types.size(), "polymer types exist");
        const PolymerType &pt = s.polymer_types[p.polymer_type];
        need(p.bead_ids.size() == pt.bead_types.size(), "a polymer has its type's bead count");
        for (size_t k = 0; k < p.bead_ids.size(); k++) {
            const uint32_t id = p.bead_ids[k];
            need(id < s.beads.size() && !bead_seen[id], "each bead belongs to one polymer");
            bead_seen[id] = 1;
            const Bead &b = s.beads[id];
            need(b.bead_type == pt.bead_types[k], "bead types follow the polymer type");
            need(b.is_monomer == (p.bead_ids.size() == 1), "is_monomer means a one-bead polymer");
            need(b.polymer_offset == k, "polymer offsets follow bead order");
        }
        for (const Bond &bd : pt.bonds) {
            const Bead &h = s.beads[p.bead_ids[bd.bead_offset_head]], &t = s.beads[p.bead_ids[bd.bead_offset_tail]];
            double r2 = 0;
            for (int d = 0; d < 3; d++) {
                double dx = h.x[d] - t.x[d];
                if (std::fabs(dx + s.box[d]) < std::fabs(dx)) dx += s.box[d];
                if (std::fabs(dx - s.box[d]) < std::fabs(dx)) dx -= s.box[d];
                r2 += dx * dx;
            }
            const double r = std::sqrt(r2);
            if (!(r < max_r)) {
                std::ostringstream m;
                m << "bond length " << r << " is not below max_r = " << max_r << " in polymer " << p.polymer_id
                  << " at beads (" << h.bead_id << "," << t.bead_id << ")";
                fail(m.str());
            }
        }
    }

    const double max_v = 0.8 / s.dt;                                // at most 0.8 of a cell per step
    const double max_f = std::min(max_v / s.dt, 1.6 / (s.dt * s.dt));
    std::unordered_set<uint32_t> hashes;
    for (const Bead &b : s.beads) {
        for (int d = 0; d < 3; d++) {
            need(0 <= b.x[d] && b.x[d] < s.box[d], "0 <= x < box");
            need(-max_v <= b.v[d] && b.v[d] <= max_v, "|v| <= 0.8/dt per component");
            need(-max_f <= b.f[d] && b.f[d] <= max_f, "|f| within the per-step limit");
        }
        need(hashes.insert(b.get_hash_code().hash).second, "bead hashes are unique");
    }
}

}  // namespace pdpd_io
