// Tool plugins: callbacks around every forall and launch
// (include/RAJA/util/PluginStrategy.hpp, PluginContext.hpp, Registry.hpp,
// plugins.hpp and src/RuntimePluginLoader.cpp:94-127 in the reference).
//
//   * a plugin derives from RAJA::util::PluginStrategy and overrides any of
//     init / preCapture / postCapture / preLaunch / postLaunch / finalize;
//   * static registration: static RAJA::util::PluginRegistry::add<MyPlugin> P("name", "desc");
//   * runtime loading: RAJA_PLUGINS names a plugin library (*.so) or a
//     directory of them; each library exports
//       extern "C" RAJA::util::PluginStrategy* RAJAGetPlugin();
//     and is loaded (dlopen) before the first forall/launch, or by
//     RAJA::util::init_plugins(path);
//   * the context carries the platform and the kernel name given with
//     RAJA::Name("...") (empty when none).
// With no plugin registered the call points cost one check of an empty list.
#pragma once

#include <dirent.h>
#include <dlfcn.h>
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "RAJA/config.hpp"
#include "RAJA/policy/PolicyBase.hpp"
#include "RAJA/util/resource.hpp"

namespace RAJA
{
namespace util
{

struct PluginContext
{
  PluginContext(Platform p, std::string name) : platform(p), kernel_name(std::move(name)) {}
  Platform platform;
  std::string kernel_name;
};

struct PluginOptions
{
  explicit PluginOptions(std::string s = "") : str(std::move(s)) {}
  std::string str;
};

inline PluginOptions make_options(const std::string& s) { return PluginOptions(s); }

class PluginStrategy
{
public:
  PluginStrategy() = default;
  virtual ~PluginStrategy() = default;
  virtual void init(const PluginOptions&) {}
  virtual void preCapture(const PluginContext&) {}
  virtual void postCapture(const PluginContext&) {}
  virtual void preLaunch(const PluginContext&) {}
  virtual void postLaunch(const PluginContext&) {}
  virtual void finalize() {}
};

namespace detail
{

struct PluginEntry
{
  std::string name, desc;
  std::unique_ptr<PluginStrategy> plugin;
};

inline std::vector<PluginEntry>& plugin_list()
{
  static std::vector<PluginEntry> list;
  return list;
}

inline bool ends_with_so(const std::string& f)
{
  return f.size() > 3 && f.compare(f.size() - 3, 3, ".so") == 0;
}

// dlopen one plugin library and register what its RAJAGetPlugin returns.
inline void load_plugin_library(const std::string& path)
{
  (void)dlerror();
  void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_GLOBAL);
  if (!h)
  {
    std::fprintf(stderr, "[RAJA plugins] dlopen failed: %s\n", dlerror());
    return;
  }
  using getter_t = PluginStrategy* (*)();
  getter_t get = reinterpret_cast<getter_t>(dlsym(h, "RAJAGetPlugin"));
  if (!get) get = reinterpret_cast<getter_t>(dlsym(h, "getPlugin"));
  if (!get)
  {
    std::fprintf(stderr, "[RAJA plugins] %s exports no RAJAGetPlugin\n", path.c_str());
    return;
  }
  plugin_list().push_back(PluginEntry{path, "runtime plugin", std::unique_ptr<PluginStrategy>(get())});
}

// A plugin library, or every *.so in a directory.
inline void load_plugin_path(const std::string& path)
{
  struct stat st;
  if (path.empty() || stat(path.c_str(), &st) != 0) return;
  if (S_ISDIR(st.st_mode))
  {
    DIR* d = opendir(path.c_str());
    if (!d) return;
    std::vector<std::string> files;
    while (dirent* e = readdir(d))
    {
      const std::string f = e->d_name;
      if (ends_with_so(f)) files.push_back(path + "/" + f);
    }
    closedir(d);
    for (const auto& f : files) load_plugin_library(f);
  }
  else if (ends_with_so(path))
  {
    load_plugin_library(path);
  }
}

// RAJA_PLUGINS is read once, before the first call point runs.
inline void load_env_plugins_once()
{
  static const bool loaded = [] {
    if (const char* env = std::getenv("RAJA_PLUGINS")) load_plugin_path(env);
    return true;
  }();
  (void)loaded;
}

}  // namespace detail

// Static registration (the reference's Registry<PluginStrategy>::add).
struct PluginRegistry
{
  template<typename P>
  struct add
  {
    add(const std::string& name, const std::string& desc)
    {
      detail::plugin_list().push_back(detail::PluginEntry{name, desc, std::unique_ptr<PluginStrategy>(new P())});
    }
  };
  static size_t size() { return detail::plugin_list().size(); }
};

// Platform a policy runs on (a segment-iteration pair: its segment policy).
namespace detail
{
template<typename P>
constexpr Platform platform_from(int, decltype(P::policy)* = nullptr)
{
  return P::policy == Policy::cdna4 ? Platform::cdna4 : Platform::host;
}
template<typename P>
constexpr Platform platform_from(long, typename P::seg_exec* = nullptr)
{
  return platform_from<typename P::seg_exec>(0);
}
template<typename P>
constexpr Platform platform_from(...)
{
  return Platform::undefined;
}
}  // namespace detail

template<typename P>
struct platform_of
{
  static constexpr Platform value = detail::platform_from<P>(0);
};

template<typename Policy>
inline PluginContext make_context(std::string name)
{
  return PluginContext(platform_of<Policy>::value, std::move(name));
}

inline void callPreCapturePlugins(const PluginContext& p)
{
  detail::load_env_plugins_once();
  for (auto& e : detail::plugin_list()) e.plugin->preCapture(p);
}
inline void callPostCapturePlugins(const PluginContext& p)
{
  for (auto& e : detail::plugin_list()) e.plugin->postCapture(p);
}
inline void callPreLaunchPlugins(const PluginContext& p)
{
  for (auto& e : detail::plugin_list()) e.plugin->preLaunch(p);
}
inline void callPostLaunchPlugins(const PluginContext& p)
{
  for (auto& e : detail::plugin_list()) e.plugin->postLaunch(p);
}
inline void callInitPlugins(const PluginOptions& p)
{
  detail::load_env_plugins_once();
  for (auto& e : detail::plugin_list()) e.plugin->init(p);
}
// Loads the plugins found at `path` (library or directory), then runs init.
inline void init_plugins(const std::string& path)
{
  detail::load_env_plugins_once();
  detail::load_plugin_path(path);
  callInitPlugins(make_options(path));
}
inline void init_plugins() { callInitPlugins(make_options("")); }
inline void finalize_plugins()
{
  for (auto& e : detail::plugin_list()) e.plugin->finalize();
}

// Brackets one kernel: pre/post capture around the body copy, pre/post
// launch around the launch.
struct PluginScope
{
  PluginContext ctx;
  explicit PluginScope(PluginContext c) : ctx(std::move(c))
  {
    callPreCapturePlugins(ctx);
    callPostCapturePlugins(ctx);
    callPreLaunchPlugins(ctx);
  }
  ~PluginScope() { callPostLaunchPlugins(ctx); }
  PluginScope(const PluginScope&) = delete;
  PluginScope& operator=(const PluginScope&) = delete;
};

}  // namespace util
}  // namespace RAJA
