// force_logging.hpp -- the reference's force-logging interface (ForceLogging,
// include/dpd/core/logging.hpp) and its CSV file logger (FileLogger,
// include/dpd/core/logging_impl.hpp), restated for tools built without the
// reference tree. With PDPD_HOST_REFERENCE the engine header uses the
// reference's own classes instead; the call sequence and the file format are
// the same, so logs from either engine can be compared line for line.
//
// File format, one property per line:
//   <prefix>Prop,<t>,<bead hash>,<bead hash>,<bead hash>,<name>,<x0>,<x1>,<x2>
// Bead hashes are printed without their type bits (BeadHash::reduced_hash);
// unused hash and value fields are empty; a value that is an integer in
// [0, 2^32) is printed as one, anything else in C++'s default format.
#pragma once

#include "pdpd/world_state.hpp"

#include <cassert>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

class ForceLogging {
public:
    virtual ~ForceLogging() {}

    virtual void Flush() = 0;
    virtual void SetTime(long t) = 0;
    virtual void SetPrefix(const std::string &) = 0;
    virtual void LogProperty(const char *name, int dims, const double *x) = 0;
    virtual void LogBeadProperty(BeadHash b, const char *name, int dims, const double *x) = 0;
    virtual void LogBeadPairProperty(BeadHash b0, BeadHash b1, const char *name, int dims, const double *x) = 0;
    virtual void LogBeadTripleProperty(BeadHash b0, BeadHash b1, BeadHash b2, const char *name, int dims,
                                       const double *x) = 0;

    // The process-wide logger; null (the default) disables logging.
    static ForceLogging *&logger()
    {
        static ForceLogging *l = nullptr;
        return l;
    }
    static void set_logger(ForceLogging *l) { logger() = l; }
};

class FileLogger : public ForceLogging {
public:
    explicit FileLogger(const std::string &name) : m_dst(name)
    {
        if (!m_dst.is_open()) fprintf(stderr, "Couldnt open force logging file %s\n", name.c_str());
    }
    ~FileLogger() override { m_dst.flush(); }

    void Flush() override { m_dst.flush(); }
    void SetTime(long t) override { m_t = t; }
    void SetPrefix(const std::string &prefix) override { m_prefix = prefix; }

    void LogProperty(const char *name, int dims, const double *x) override { line(nullptr, 0, name, dims, x); }
    void LogBeadProperty(BeadHash b, const char *name, int dims, const double *x) override
    {
        const BeadHash h[1] = {b};
        line(h, 1, name, dims, x);
    }
    void LogBeadPairProperty(BeadHash b0, BeadHash b1, const char *name, int dims, const double *x) override
    {
        const BeadHash h[2] = {b0, b1};
        line(h, 2, name, dims, x);
    }
    void LogBeadTripleProperty(BeadHash b0, BeadHash b1, BeadHash b2, const char *name, int dims,
                               const double *x) override
    {
        const BeadHash h[3] = {b0, b1, b2};
        line(h, 3, name, dims, x);
    }

private:
    void value(double v)
    {
        if (std::round(v) == v && 0 <= v && v < 4294967296.0) m_dst << (uint32_t)v;
        else m_dst << v;
    }

    void line(const BeadHash *h, int nh, const char *name, int dims, const double *x)
    {
        assert(dims <= 3);
        m_dst << m_prefix << "Prop," << m_t;
        for (int k = 0; k < 3; k++) {
            m_dst << ",";
            if (k < nh) m_dst << h[k].reduced_hash().hash;
        }
        m_dst << "," << name;
        for (int k = 0; k < 3; k++) {
            m_dst << ",";
            if (k < dims) value(x[k]);
        }
        m_dst << "\n";
    }

    std::ofstream m_dst;
    long m_t = 0;
    std::string m_prefix;
};
