// dpd_engine.hpp -- the DPDEngine plugin interface and name registry of
// poets-dpd (include/dpd/core/dpd_engine.hpp:10-261), restated here so the
// repository's engines and tools build without the reference tree.
//
// The virtual functions, their signatures and their contracts are the
// reference's: CanSupport returns "" or a reason; Attach(nullptr) detaches;
// Run(n) leaves the WorldState synchronised with t advanced by n; the interval
// Run calls the callback after each interval and stops when it returns false;
// the factory throws on duplicate or unknown names.
#pragma once

#include "pdpd/world_state.hpp"

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

class DPDEngine {
protected:
    virtual bool CanSupportHookeanBonds() const { return false; }
    virtual bool CanSupportAngleBonds() const { return false; }
    virtual bool CanSupportStationaryBeadTypes() const { return false; }
    virtual bool CanSupportSpatialDPDParameters() const { return false; }
    virtual bool CanSupportGeneralDPDParameters() const { return false; }
    virtual bool CanSupportSpatialBondParameters() const { return false; }
    virtual bool CanSupportGeneralBondParameters() const { return false; }

public:
    virtual ~DPDEngine() {}

    struct timings_t {
        double compile = -1;
        double aquire = -1;
        double configure = -1;
        double execute_to_first_bead = -1;
        double execute_to_last_bead = -1;
        double perf_counters = -1;
    };

    // Capability checks in the reference's order: parameter expressions, then
    // stationary types, then bond kinds (dpd_engine.hpp:57-128).
    virtual std::string CanSupport(const WorldState *s) const
    {
        const std::unordered_set<std::string> space{"x", "y", "z", "ux", "uy", "uz"};
        auto classify = [&](const std::unordered_set<std::string> &vars, bool spatial_ok, bool general_ok,
                            const char *what) -> std::string {
            if (vars.empty()) return "";
            bool space_only = true;
            for (const auto &k : vars) space_only &= space.count(k) > 0;
            if (space_only) {
                if (!spatial_ok) return std::string("Engine does not support space varying ") + what + " parameters.";
            } else if (!general_ok) {
                return std::string("Engine does not support aribtrary varying ") + what + " parameters.";
            }
            return "";
        };
        std::unordered_set<std::string> ivars;
        for (const auto &ii : s->interactions) {
            ii.conservative.collect_variables(ivars);
            ii.dissipative.collect_variables(ivars);
        }
        std::string r = classify(ivars, CanSupportSpatialDPDParameters(), CanSupportGeneralDPDParameters(), "DPD");
        if (!r.empty()) return r;
        for (const auto &b : s->bead_types)
            if (b.stationary && !CanSupportStationaryBeadTypes()) return "Engine does not support stationary bead types.";
        std::unordered_set<std::string> bvars;
        for (const auto &p : s->polymer_types) {
            if (!p.bonds.empty() && !CanSupportHookeanBonds()) return "Engine does not support hookean bonds.";
            if (!p.bond_pairs.empty() && !CanSupportAngleBonds()) return "Engine does not support angle bonds.";
            for (const auto &b : p.bonds) { b.kappa.collect_variables(bvars); b.r0.collect_variables(bvars); }
            for (const auto &bp : p.bond_pairs) { bp.kappa.collect_variables(bvars); bp.theta0.collect_variables(bvars); }
        }
        return classify(bvars, CanSupportSpatialBondParameters(), CanSupportGeneralBondParameters(), "bond");
    }

    virtual double GetMaxBondLength() const { return 1.0; }
    virtual void PrepareResources() {}
    virtual bool GetTimings(timings_t &) { return false; }

    struct engine_config_entry_t {
        std::string name;
        std::string description;
        std::string value;
    };

    virtual void GetConfigEntries(std::vector<engine_config_entry_t> &config) { config.clear(); }

    virtual void SetConfigEntry(const std::string &key, const std::string &)
    {
        throw std::runtime_error("Unknown engine config name '" + key + "'");
    }

    virtual void Attach(WorldState *state) = 0;
    virtual void Run(unsigned nSteps) = 0;

    virtual unsigned Run(int interval_count, unsigned interval_size, std::function<bool()> interval_callback)
    {
        unsigned done = 0;
        while (interval_count != 0) {
            Run(interval_size);
            done += interval_size;
            if (!interval_callback()) break;
            if (interval_count > 0) --interval_count;
        }
        return done;
    }
};

class DPDEngineFactory {
public:
    using creator_func_t = std::function<std::shared_ptr<DPDEngine>()>;

    static std::vector<std::string> ListFactories()
    {
        std::vector<std::string> names;
        for (const auto &kv : registry()) names.push_back(kv.first);
        return names;
    }

    static bool RegisterFactory(const std::string &name, creator_func_t func)
    {
        if (!registry().emplace(name, std::move(func)).second)
            throw std::runtime_error("Duplicate DPD engines registered as '" + name + "'");
        return true;
    }

    static std::shared_ptr<DPDEngine> create(const std::string &name)
    {
        auto it = registry().find(name);
        if (it == registry().end()) throw std::runtime_error("No DPD engine registered called '" + name + "'");
        return it->second();
    }

private:
    static std::map<std::string, creator_func_t> &registry()
    {
        static std::map<std::string, creator_func_t> m;
        return m;
    }
};
