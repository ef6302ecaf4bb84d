#pragma once

// =============================================================================
// cospanmdao_amd/workloads/ns_systems.hpp — the NS residual as a System: fused (one fused group) or unfused
// =============================================================================
//
// Ports: inputs u v p (the stored fields), coef ({1 / 2h, 1 / h^2, nu}); outputs Ru Rv Rc, sq (FP64 sums of squares),
// norm (FP64 L2 norms of Ru, Rv, Rc and of all three).
//
//   fused    fuse_components("ns", connects, NsTile<S>, global_barrier(), NsNorm): one fused group (the structure of the
//            reference's fused ns-residual: tile stage, barrier, deterministic reduction; examples/ns-residual/systems.hpp:88-140)
//   unfused  NsGradients -> NsCombine -> NsSquares -> NsNorm: four launches, derivatives through global memory

#include <cstdint>
#include <functional>
#include <string>

#include "naga/cospan/cospan.hpp"
#include "ns_components.hpp"

namespace ns {

enum class Style { fused = 0, unfused = 1 };

/// The field initializer (defined by the library that instantiates the System: it owns the combine kernel).
using FieldInit = std::function<cs::InitSpec(int field)>;

template <DType S>
class NsResidual : public cs::System {
public:
    NsResidual(Style style, int nx, int ny, double h, double nu, int ty, FieldInit fields)
        : style_{style}, nx_{nx}, ny_{ny}, h_{h}, nu_{nu}, ty_{ty}, fields_{std::move(fields)} {}

    void setup() override {
        if (style_ == Style::fused) {
            add_subsystem(cs::fuse_components("ns", cs::connects_desc({{"tile.part", "norm.part"}}), cs::subsystem_desc<NsTile<S>>("tile", ty_),
                                              cs::global_barrier(), cs::subsystem_desc<NsNorm>("norm")));
            promote<S, 2>("ns.tile.u", "u");
            promote<S, 2>("ns.tile.v", "v");
            promote<S, 2>("ns.tile.p", "p");
            promote<DType::f32, 1>("ns.tile.coef", "coef");
            promote<DType::f32, 2>("ns.tile.Ru", "Ru");
            promote<DType::f32, 2>("ns.tile.Rv", "Rv");
            promote<DType::f32, 2>("ns.tile.Rc", "Rc");
            promote<DType::f64, 1>("ns.norm.sq", "sq");
            promote<DType::f64, 1>("ns.norm.norm", "norm");
        } else {
            add_subsystem(cs::subsystem_desc<NsGradients<S>>("grad"));
            add_subsystem(cs::subsystem_desc<NsCombine<S>>("comb"));
            add_subsystem(cs::subsystem_desc<NsSquares>("squares"));
            add_subsystem(cs::subsystem_desc<NsNorm>("norm"));
            connect("grad.d", "comb.d");
            connect("comb.Ru", "squares.Ru");
            connect("comb.Rv", "squares.Rv");
            connect("comb.Rc", "squares.Rc");
            connect("squares.part", "norm.part");
            tie_inputs<S, 2>({"grad.u", "comb.u"}, "u");
            tie_inputs<S, 2>({"grad.v", "comb.v"}, "v");
            tie_inputs<DType::f32, 1>({"grad.coef", "comb.coef"}, "coef");
            promote<S, 2>("grad.p", "p");
            promote<DType::f32, 2>("comb.Ru", "Ru");
            promote<DType::f32, 2>("comb.Rv", "Rv");
            promote<DType::f32, 2>("comb.Rc", "Rc");
            promote<DType::f64, 1>("norm.sq", "sq");
            promote<DType::f64, 1>("norm.norm", "norm");
        }
        init_input("u", fields_(0));
        init_input("v", fields_(1));
        init_input("p", fields_(2));
        const float inv2h = static_cast<float>(1.0 / (2.0 * h_)), invh2 = static_cast<float>(1.0 / (h_ * h_));
        init_input("coef", cs::init::from_host(std::vector<float>{inv2h, invh2, static_cast<float>(nu_)}));
    }

private:
    Style     style_;
    int       nx_, ny_;
    double    h_, nu_;
    int       ty_;
    FieldInit fields_;
};

}  // namespace ns
