// Runtime tool plugin: one roctx range per RAJA forall / launch, named by
// RAJA::Name (or "RAJA kernel"), so rocprofv3 --marker-trace shows the
// loops of an unchanged RAJA program.  The reference builds the same
// ranges into its HIP backend when RAJA_ENABLE_ROCTX is on
// (include/RAJA/policy/hip/forall.hpp); here they come from a plugin:
//   RAJA_PLUGINS=<repo>/raja_amd/plugins/libroctx_plugin.so ./app
#include <rocprofiler-sdk-roctx/roctx.h>

#include "RAJA/util/plugins.hpp"

namespace
{

class RoctxPlugin : public RAJA::util::PluginStrategy
{
public:
  void preLaunch(const RAJA::util::PluginContext& p) override
  {
    roctxRangePush(p.kernel_name.empty() ? "RAJA kernel" : p.kernel_name.c_str());
  }
  void postLaunch(const RAJA::util::PluginContext&) override { roctxRangePop(); }
};

}  // namespace

extern "C" RAJA::util::PluginStrategy* RAJAGetPlugin() { return new RoctxPlugin(); }
