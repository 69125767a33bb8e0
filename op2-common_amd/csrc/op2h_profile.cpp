// op2h_profile.cpp — the timing tree behind the op_profile_* C API.
//
// Behaviour follows the reference's op_profile (op2/include/op_profile.h:15-38,
// 197-212; op2/src/externlib/op_profile.cpp:214-572), written for one rank and
// one HIP stream:
//   - OP_PROFILE_LEVEL, read at op_profile_start: 0 off; 1 (the default, as
//     op_profile.h:131 initialises it) the application's own sections; 2 adds
//     one node per op_par_loop; 3 adds, under each loop node, "Computation"
//     (launch to completion, wall clock) holding "Kernel" (the kernels' own
//     device time from HIP events);
//   - sections nest by name; a section entered below a loop node at level < 3
//     is not timed (op_profile.cpp:259-265);
//   - each node keeps count, total, min and max;
//   - op_profile_output prints the tree of non-loop sections, then a table of
//     loop nodes per section (sorted by total time, at most 20 rows, "%kern"
//     at level 3), and writes JSON to OP_PROFILE_JSON_OUTPUT when it is set;
//     op_profile_output_json writes JSON to a named file.  The JSON has the
//     reference's shape: {"name", "type" (0 section, 1 kernel), "clock":
//     {"n", "total", "min", "max"} in nanoseconds, "num_ranks", "children"}.
// The device is synchronised at section boundaries, as the reference does,
// except while a HIP graph is being captured (a captured step has no host
// boundaries: sections entered then record host time only and loop nodes are
// not created).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "op2h_profile.h"

extern "C" void op2h_device_sync(void);

namespace {

using ns_t = long long;

ns_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Clock {
  long long n = 0;
  ns_t total = 0, min = 0, max = 0;
  void submit(ns_t d) {
    if (n == 0) {
      total = min = max = d;
    } else {
      total += d;
      min = std::min(min, d);
      max = std::max(max, d);
    }
    ++n;
  }
};

struct Node {
  std::string name;
  int kind = 0;  // 0 section, 1 loop (kernel)
  Clock clock;
  std::vector<std::unique_ptr<Node>> children;
  explicit Node(std::string n, int k = 0) : name(std::move(n)), kind(k) {}
  Node &child(const std::string &n, int k = 0) {
    for (auto &c : children)
      if (c->name == n) return *c;
    children.emplace_back(new Node(n, k));
    return *children.back();
  }
  const Node *find(const std::string &n) const {
    for (auto &c : children)
      if (c->name == n) return c.get();
    return nullptr;
  }
};

struct Profile {
  int level = 1;
  bool started = false, finished = false;
  std::unique_ptr<Node> root;
  std::vector<Node *> scope;
  std::vector<ns_t> starts;
  int untimed_depth = 0;  // sections entered below a level's detail
};

Profile &prof() {
  static Profile P;
  return P;
}

void sync_device() { op2h_device_sync(); }  // a no-op while capturing

void enter_section(const std::string &name, bool sync) {
  Profile &P = prof();
  if (P.level == 0 || !P.started || P.finished) return;
  if (P.untimed_depth > 0 || (P.level < 3 && P.scope.back()->kind == 1)) {
    ++P.untimed_depth;
    return;
  }
  Node &n = P.scope.back()->child(name);
  if (sync) sync_device();
  roctxRangePushA(name.c_str());  // a rocprofv3 --marker-trace range
  P.scope.push_back(&n);
  P.starts.push_back(now_ns());
}

void exit_section(bool sync) {
  Profile &P = prof();
  if (P.level == 0 || !P.started || P.finished) return;
  if (P.untimed_depth > 0) {
    --P.untimed_depth;
    return;
  }
  if (P.scope.size() <= 1) return;  // the root closes at op_profile_end
  if (sync) sync_device();
  P.scope.back()->clock.submit(now_ns() - P.starts.back());
  roctxRangePop();
  P.scope.pop_back();
  P.starts.pop_back();
}

// Duration printed in `width` characters with a unit picked so that at least
// one digit stands left of the point (s, ms, us, ns).
std::string fmt_duration(ns_t d, int width = 7) {
  static const char *unit[4] = {"s", "ms", "us", "ns"};
  double v = d * 1e-9;
  int u = 0;
  while (u < 3 && v < 1.0) {
    v *= 1000.0;
    ++u;
  }
  const int uw = (int)std::strlen(unit[u]);
  int left = 1;
  for (double t = v; t >= 10.0; t /= 10.0) ++left;
  const int room = width - uw;
  const int decimals = std::max(room - left - 1, 0);
  char buf[64];
  if (decimals > 0)
    std::snprintf(buf, sizeof buf, "%*.*f%s", room, decimals, v, unit[u]);
  else
    std::snprintf(buf, sizeof buf, "%*.0f%s", room, v, unit[u]);
  return buf;
}

std::string clock_line(const Clock &c, const Clock *parent) {
  std::string s = "*" + std::to_string(c.n) + " total: " + fmt_duration(c.total);
  if (parent && parent->total > 0) {
    char b[32];
    std::snprintf(b, sizeof b, " (%.1f%%)", 100.0 * (double)c.total / (double)parent->total);
    s += b;
  }
  if (c.n > 1)
    s += " (average: " + fmt_duration(c.total / c.n) + ", min: " + fmt_duration(c.min) +
         ", max: " + fmt_duration(c.max) + ")";
  return s;
}

void print_tree(const Node &n, int indent, const Clock *parent) {
  std::printf("%*s%s %s\n", indent, "", n.name.c_str(), clock_line(n.clock, parent).c_str());
  for (auto &c : n.children) print_tree(*c, indent + 4, parent ? parent : &n.clock);
}

void walk_sections(const Node &n, int indent, std::vector<const Node *> &with_loops,
                   std::vector<std::string> &paths, const std::string &path) {
  bool has_loop = false;
  for (auto &c : n.children) has_loop |= c->kind == 1;
  const std::string here = path.empty() ? n.name : path + "/" + n.name;
  if (has_loop) {
    with_loops.push_back(&n);
    paths.push_back(here);
  }
  std::printf("%*s%s %s\n", indent, "", n.name.c_str(), clock_line(n.clock, nullptr).c_str());
  for (auto &c : n.children)
    if (c->kind == 0) walk_sections(*c, indent + 4, with_loops, paths, here);
}

void print_loops(const Node &n, const std::string &path, size_t width, int level) {
  std::printf("%s%*s     num    total      avg      min      max%s\n", path.c_str(),
              (int)(width + 4 - path.size()), "", level >= 3 ? "    %kern" : "");
  std::vector<const Node *> loops;
  for (auto &c : n.children)
    if (c->kind == 1) loops.push_back(c.get());
  std::sort(loops.begin(), loops.end(),
            [](const Node *a, const Node *b) { return a->clock.total > b->clock.total; });
  const size_t limit = 20;
  for (size_t i = 0; i < loops.size(); ++i) {
    if (i == limit) {
      std::printf("    (%zu more)\n", loops.size() - limit);
      break;
    }
    const Node &l = *loops[i];
    std::string kern;
    if (level >= 3) {
      const Node *comp = l.find("Computation");
      const Node *k = comp ? comp->find("Kernel") : nullptr;
      const Node *use = k ? k : comp;
      char b[16];
      std::snprintf(b, sizeof b, "    %5.1f",
                    use && l.clock.total > 0 ? 100.0 * use->clock.total / l.clock.total : 0.0);
      kern = b;
    }
    std::printf("    %s%*s    %8lld  %s  %s  %s  %s%s\n", l.name.c_str(),
                (int)(width - std::min(width, l.name.size() + 4)), "", l.clock.n,
                fmt_duration(l.clock.total).c_str(),
                fmt_duration(l.clock.n ? l.clock.total / l.clock.n : 0).c_str(),
                fmt_duration(l.clock.min).c_str(), fmt_duration(l.clock.max).c_str(),
                kern.c_str());
  }
  std::printf("\n");
  if (level < 3) return;
  for (size_t i = 0; i < loops.size() && i < 4; ++i) {
    print_tree(*loops[i], 4, nullptr);
    std::printf("\n");
  }
}

void json_string(FILE *f, const std::string &s) {
  std::fputc('"', f);
  for (unsigned char ch : s) {
    if (ch == '"' || ch == '\\')
      std::fprintf(f, "\\%c", ch);
    else if (ch < 0x20)
      std::fprintf(f, "\\u%04x", ch);
    else
      std::fputc(ch, f);
  }
  std::fputc('"', f);
}

void json_node(FILE *f, const Node &n) {
  std::fprintf(f, "{\"name\":");
  json_string(f, n.name);
  std::fprintf(f, ",\"type\":%d,\"clock\":{\"n\":%lld,\"total\":%lld,\"min\":%lld,\"max\":%lld}"
                  ",\"num_ranks\":1,\"children\":[",
               n.kind, n.clock.n, n.clock.total, n.clock.min, n.clock.max);
  for (size_t i = 0; i < n.children.size(); ++i) {
    if (i) std::fputc(',', f);
    json_node(f, *n.children[i]);
  }
  std::fprintf(f, "]}");
}

}  // namespace

namespace op2h {

bool profile_loops_on() {
  const Profile &P = prof();
  return P.started && !P.finished && P.level >= 2 && P.untimed_depth == 0;
}

int profile_level() { return prof().level; }

void profile_loop(const char *name, const char *variant, long long wall_ns,
                  long long comp_ns, double kernel_ms) {
  Profile &P = prof();
  if (!profile_loops_on()) return;
  std::string full = std::string(name) + "/hip/" + variant;
  Node &l = P.scope.back()->child(full, 1);
  l.kind = 1;
  l.clock.submit(wall_ns);
  if (P.level >= 3) {
    Node &c = l.child("Computation");
    c.clock.submit(comp_ns);
    if (kernel_ms >= 0.0) c.child("Kernel").clock.submit((ns_t)(kernel_ms * 1e6));
  }
}

}  // namespace op2h

extern "C" void op_profile_start(const char *name) {
  Profile &P = prof();
  if (P.started) {
    std::printf("op_profile_start error -- timing already started\n");
    exit(-1);
  }
  if (const char *s = std::getenv("OP_PROFILE_LEVEL")) {
    char *end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (end == s || *end != '\0' || v < 0 || v > 3)
      std::printf("warning: OP_PROFILE_LEVEL set to unsupported value: %s\n", s);
    else
      P.level = (int)v;
  }
  if (P.level == 0) return;
  P.started = true;
  sync_device();
  P.root.reset(new Node(name));
  P.scope.assign(1, P.root.get());
  P.starts.assign(1, now_ns());
}

extern "C" void op_profile_enter(const char *name) { enter_section(name, true); }

extern "C" void op_profile_enter_kernel(const char *name, const char *target,
                                        const char *variant) {
  Profile &P = prof();
  if (P.level == 0 || !P.started || P.finished) return;
  if (P.level < 2) {
    ++P.untimed_depth;
    return;
  }
  enter_section(std::string(name) + "/" + target + "/" + variant, true);
  if (P.untimed_depth == 0) P.scope.back()->kind = 1;
}

extern "C" void op_profile_next(const char *name) {
  exit_section(true);
  enter_section(name, false);
}

extern "C" void op_profile_exit(void) { exit_section(true); }

extern "C" void op_profile_end(void) {
  Profile &P = prof();
  if (P.level == 0 || !P.started || P.finished) return;
  sync_device();
  const ns_t t = now_ns();
  P.untimed_depth = 0;
  while (!P.scope.empty()) {
    P.scope.back()->clock.submit(t - P.starts.back());
    if (P.scope.size() > 1) roctxRangePop();
    P.scope.pop_back();
    P.starts.pop_back();
  }
  P.finished = true;
}

extern "C" void op_profile_output_json(const char *filename) {
  Profile &P = prof();
  if (P.level == 0 || !P.root) return;
  if (!P.finished) op_profile_end();
  FILE *f = std::fopen(filename, "w");
  if (!f) {
    std::printf("Unable to open timing JSON output file: %s (%s)\n", filename,
                std::strerror(errno));
    return;
  }
  std::printf("Writing timing JSON to: %s\n", filename);
  json_node(f, *P.root);
  std::fclose(f);
}

extern "C" void op_profile_output(void) {
  Profile &P = prof();
  if (P.level == 0 || !P.root) return;
  if (!P.finished) op_profile_end();
  std::vector<const Node *> with_loops;
  std::vector<std::string> paths;
  walk_sections(*P.root, 0, with_loops, paths, "");
  std::printf("\n");
  size_t width = 0;
  for (size_t i = 0; i < with_loops.size(); ++i) {
    width = std::max(width, paths[i].size() + 1);
    for (auto &c : with_loops[i]->children) width = std::max(width, c->name.size() + 4);
  }
  for (size_t i = 0; i < with_loops.size(); ++i)
    print_loops(*with_loops[i], paths[i], width, P.level);
  if (const char *j = std::getenv("OP_PROFILE_JSON_OUTPUT")) op_profile_output_json(j);
}
