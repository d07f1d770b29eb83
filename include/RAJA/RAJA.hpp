// Umbrella header of the cdna4 RAJA front-end subset: forall, launch,
// reducers (old style, multi (binned) and expt::Reduce parameters), scans,
// sorts, WorkGroup,
// resources and synchronize, for the seq_exec, omp_parallel_for_exec and
// cdna4 (MI355X) policy families.  The cdna4 family is enabled when the
// translation unit is compiled by hipcc.  Backends are included before the
// scan/sort front-ends, whose calls into RAJA::impl are qualified names.
#pragma once

#include "RAJA/config.hpp"
#include "RAJA/util/Operators.hpp"
#include "RAJA/util/Span.hpp"
#include "RAJA/util/resource.hpp"
#include "RAJA/index/RangeSegment.hpp"
#include "RAJA/index/ListSegment.hpp"
#include "RAJA/index/IndexSet.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/pattern/params.hpp"
#include "RAJA/pattern/reduce.hpp"
#include "RAJA/pattern/multi_reduce.hpp"
#include "RAJA/policy/host.hpp"
#include "RAJA/pattern/WorkGroup.hpp"
#include "RAJA/pattern/launch.hpp"
#if defined(RAJA_ENABLE_CDNA4)
#include "RAJA/policy/cdna4.hpp"
#endif
#include "RAJA/pattern/forall.hpp"
#include "RAJA/pattern/scan.hpp"
#include "RAJA/pattern/sort.hpp"
#include "RAJA/pattern/synchronize.hpp"
