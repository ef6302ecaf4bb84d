#pragma once

// =============================================================================
// cospanmdao_amd/workloads/nsc_systems.hpp — the compressible NS residual as Systems (the reference's
// examples/ns-residual/systems.hpp:26-140 variants, on this build's components)
// =============================================================================
//
// Every variant exposes the same paths:
//   input   U                    the conserved state [4, ny, nx]
//   output  R                    the residual [4, ny, nx]
//           Dt Ws Sq             the ATOMIC diagnostics (tied: reset by init, accumulated by the tile, read by err)
//           DetDt DetWs DetSq    the deterministic ones
//           Err                  atomic - deterministic: [dt, ws, sq0 .. sq3]
//           Part                 the per-tile partials [6, tiles]
//
//   Variant::unfused   every stage its own launch: init, tile, detn, detd, detw, err (ties by System::tie_outputs)
//   Variant::fused     ONE launch with group waits, the reference's layout (systems.hpp:88-140):
//                        0 init (1 group; no barrier after it: the tile's atomics wait on it)
//                        1 tile (one group per tile)
//                        global_barrier()
//                        2 deterministic reductions (range_mode::own: 4 + 1 + 1 groups)
//                        3 err (waits on exactly the groups of segment 2)
//   Variant::launches  the same group with global barriers after init, the tile and the reductions, and no group waits: each
//                      barrier is a launch boundary (four launches of one kernel)

#include <stdexcept>
#include <string>

#include "naga/cospan/cospan.hpp"
#include "nsc_components.hpp"

namespace nsc {

enum class Variant { unfused = 0, fused = 1, launches = 2 };

template <int Lim, bool Walls, int TY>
class NscResidual : public cs::System {
public:
    explicit NscResidual(Variant v) : v_{v} {}

    void setup() override {
        using Tile = NscTile<Lim, Walls, TY>;
        if (v_ == Variant::unfused) {
            add_subsystem(cs::subsystem_desc<NscInit>("init"));
            add_subsystem(cs::subsystem_desc<Tile>("tile"));
            add_subsystem(cs::subsystem_desc<NscDet<0>>("detn"));
            add_subsystem(cs::subsystem_desc<NscDet<1>>("detd"));
            add_subsystem(cs::subsystem_desc<NscDet<2>>("detw"));
            add_subsystem(cs::subsystem_desc<NscError>("err"));
            for (const char* d : {"detn", "detd", "detw"}) connect("tile.Part", std::string{d} + ".Part");
            connect("detn.Det", "err.DetSq");
            connect("detd.Det", "err.DetDt");
            connect("detw.Det", "err.DetWs");
            tie_outputs<DType::atomic_f32, 1>({"init.Dt", "tile.Dt", "err.Dt"}, "Dt", cs::shape(1));
            tie_outputs<DType::atomic_f32, 1>({"init.Ws", "tile.Ws", "err.Ws"}, "Ws", cs::shape(1));
            tie_outputs<DType::atomic_f32, 1>({"init.Sq", "tile.Sq", "err.Sq"}, "Sq", cs::shape(4));
            expose("");
            return;
        }
        const auto connects = cs::connects_desc({{"tile.Part", "detn.Part"},
                                                 {"tile.Part", "detd.Part"},
                                                 {"tile.Part", "detw.Part"},
                                                 {"detn.Det", "err.DetSq"},
                                                 {"detd.Det", "err.DetDt"},
                                                 {"detw.Det", "err.DetWs"}});
        const auto ties     = cs::ties_desc({cs::tie_outputs<DType::atomic_f32, 1>({"init.Dt", "tile.Dt", "err.Dt"}, "Dt", cs::shape(1)),
                                             cs::tie_outputs<DType::atomic_f32, 1>({"init.Ws", "tile.Ws", "err.Ws"}, "Ws", cs::shape(1)),
                                             cs::tie_outputs<DType::atomic_f32, 1>({"init.Sq", "tile.Sq", "err.Sq"}, "Sq", cs::shape(4))});
        const auto dets = cs::segment<cs::range_mode::own>(cs::subsystem_desc<NscDet<0>>("detn"), cs::subsystem_desc<NscDet<1>>("detd"),
                                                           cs::subsystem_desc<NscDet<2>>("detw"));
        if (v_ == Variant::fused)
            add_subsystem(cs::fuse_components<cs::group_waits_enabled>("ns", connects, ties, cs::segment(cs::subsystem_desc<NscInit>("init")),
                                                                       cs::segment(cs::subsystem_desc<Tile>("tile")), cs::global_barrier(), dets,
                                                                       cs::segment(cs::subsystem_desc<NscError>("err"))));
        else
            add_subsystem(cs::fuse_components("ns", connects, ties, cs::segment(cs::subsystem_desc<NscInit>("init")), cs::global_barrier(),
                                              cs::segment(cs::subsystem_desc<Tile>("tile")), cs::global_barrier(), dets, cs::global_barrier(),
                                              cs::segment(cs::subsystem_desc<NscError>("err"))));
        expose("ns.");
    }

private:
    /// The uniform paths of every variant.
    void expose(const std::string& p) {
        promote<DType::f32, 3>(p + "tile.U", "U");
        promote<DType::f32, 3>(p + "tile.R", "R");
        promote<DType::f32, 2>(p + "tile.Part", "Part");
        promote<DType::f32, 1>(p + "detd.Det", "DetDt");
        promote<DType::f32, 1>(p + "detw.Det", "DetWs");
        promote<DType::f32, 1>(p + "detn.Det", "DetSq");
        promote<DType::f32, 1>(p + "err.Err", "Err");
        if (!p.empty()) {
            promote<DType::atomic_f32, 1>(p + "Dt", "Dt");
            promote<DType::atomic_f32, 1>(p + "Ws", "Ws");
            promote<DType::atomic_f32, 1>(p + "Sq", "Sq");
        }
    }
    Variant v_;
};

}  // namespace nsc
