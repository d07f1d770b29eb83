// Test plugin loaded through RAJA_PLUGINS: counts the call points it sees
// and prints them when the program finalizes the plugins.
#include <cstdio>
#include <set>
#include <string>

#include "RAJA/util/plugins.hpp"

namespace
{

class CounterPlugin : public RAJA::util::PluginStrategy
{
public:
  void init(const RAJA::util::PluginOptions&) override { ++inits; }
  void preCapture(const RAJA::util::PluginContext&) override { ++captures; }
  void preLaunch(const RAJA::util::PluginContext& p) override
  {
    ++pre;
    if (p.platform == RAJA::Platform::cdna4) ++device;
    if (!p.kernel_name.empty()) names.insert(p.kernel_name);
  }
  void postLaunch(const RAJA::util::PluginContext&) override { ++post; }
  void finalize() override
  {
    std::printf("[counter plugin] init=%ld pre=%ld post=%ld capture=%ld device=%ld names=", inits, pre,
                post, captures, device);
    for (const auto& n : names) std::printf("%s;", n.c_str());
    std::printf("\n");
  }

private:
  long inits = 0, pre = 0, post = 0, captures = 0, device = 0;
  std::set<std::string> names;
};

}  // namespace

extern "C" RAJA::util::PluginStrategy* RAJAGetPlugin() { return new CounterPlugin(); }
