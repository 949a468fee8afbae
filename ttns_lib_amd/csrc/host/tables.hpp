// Serialisation of the compiled term tables into a flat int64 stream (the layout
// documented at ttn_sop_compiled_tables in include/ttn_capi.h).
#pragma once

#include "sop.hpp"

namespace ttnh {

inline std::vector<i64> serialise_tables(const CompiledSOP& c) {
    std::vector<i64> out;
    for (const NodeOps& no : c.nodes) {
        out.push_back(no.ncommon_spf);
        out.push_back(no.ncommon_mf);
        out.push_back(no.nstandard);
        for (const Term& t : no.terms) {
            out.push_back(t.id_spf);
            out.push_back(t.id_mf);
            out.push_back(t.nspf_terms());
            for (const auto& s : t.spf_index) {
                out.push_back(static_cast<i64>(s.size()));
                for (const auto& p : s) {
                    out.push_back(p.first);
                    out.push_back(p.second);
                }
            }
            out.push_back(t.nmf_terms());
            for (const MfIndex& m : t.mf_index) {
                out.push_back(m.parent_index);
                out.push_back(static_cast<i64>(m.siblings.size()));
                for (const auto& p : m.siblings) {
                    out.push_back(p.first);
                    out.push_back(p.second);
                }
            }
        }
    }
    return out;
}

}  // namespace ttnh
