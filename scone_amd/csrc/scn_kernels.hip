// scn_kernels.hip -- explicit instantiation of one group of transport-kernel variants.
//
// Compiled once per group with
//   -DSCN_KENGINE=0|1 (0: eigenvalue engine k_transport, 1: batch engine k_btransport)
//   -DSCN_KMODE=1|2|3 (SCN_TRANSPORT_ST / _DT / _HT)
//   -DSCN_KGEO=0|1    (lattice/pin/box kernels at depths 1, 3, 4 and 8, or the general-geometry
//                      kernels, one depth per object: -DSCN_KNL=3|8)
// so the heavy kernel variants compile in parallel; scn_engine.hip and scn_batch.hip call the
// variant functions these objects define.
#if !defined(SCN_KENGINE) || !defined(SCN_KMODE) || !defined(SCN_KGEO)
#error "scn_kernels.hip needs -DSCN_KENGINE, -DSCN_KMODE and -DSCN_KGEO"
#endif
#if SCN_KGEO == 1 && !defined(SCN_KNL)
#error "general-geometry kernels are built one depth at a time (-DSCN_KNL)"
#endif

#if SCN_KENGINE == 0
// the eigenvalue kernels keep the general placement: the flat-lattice branch costs them scratch
// (an index array leaves registers) for geometries they rarely see; both give the same bits
#define SCN_FLAT_PLACEMENT 0
#define SCN_TRANSPORT_KERNELS 1
#include "scn_transport.h"
#define SCN_INST(NL) template TransportFn variant_fn<NL, SCN_KMODE, SCN_KGEO>(const void** kern);
namespace scn_eng {
#else
#define SCN_BTRANSPORT_KERNELS 1
#include "scn_btransport.h"
#define SCN_INST(NL) template void bt_variant<NL, SCN_KMODE, SCN_KGEO>(BTransportFn * fn, const void** kern, BSourceFn* src);
namespace scn_bat {
#endif

#if SCN_KGEO == 0
SCN_INST(1)
SCN_INST(3)
SCN_INST(4)
SCN_INST(8)
#else
SCN_INST(SCN_KNL)
#endif

}  // namespace
