// Minimal test harness for the C++ API programs: TEST(Suite, Name) registers a case;
// EXPECT_* record failures and continue, ASSERT_* abort the case; main() runs them all
// and returns non-zero if any failed.
#pragma once

#include <cmath>
#include <cstdio>
#include <exception>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

namespace check {

struct Case {
  std::string name;
  std::function<void()> fn;
};
inline std::vector<Case>& registry() {
  static std::vector<Case> r;
  return r;
}
inline int& failures() {
  static int n = 0;
  return n;
}
struct Abort {};
struct Reg {
  Reg(const char* s, const char* n, std::function<void()> f) {
    registry().push_back({std::string(s) + "." + n, std::move(f)});
  }
};
inline void fail(const char* file, int line, const std::string& what, bool fatal) {
  std::printf("  FAILED %s:%d: %s\n", file, line, what.c_str());
  ++failures();
  if (fatal) throw Abort{};
}
template <class A, class B>
std::string pair_str(const char* ea, const char* eb, const A& a, const B& b) {
  std::ostringstream os;
  os.precision(17);
  os << ea << " = " << a << ", " << eb << " = " << b;
  return os.str();
}

}  // namespace check

#define TEST(S, N)                                            \
  static void S##_##N##_body();                               \
  static check::Reg S##_##N##_reg(#S, #N, &S##_##N##_body);   \
  static void S##_##N##_body()

#define GC_CHECK_(cond, what, fatal) \
  do {                               \
    if (!(cond)) check::fail(__FILE__, __LINE__, what, fatal); \
  } while (0)

#define EXPECT_TRUE(c) GC_CHECK_((c), #c, false)
#define ASSERT_TRUE(c) GC_CHECK_((c), #c, true)
#define EXPECT_EQ(a, b) GC_CHECK_((a) == (b), check::pair_str(#a, #b, (a), (b)), false)
#define ASSERT_EQ(a, b) GC_CHECK_((a) == (b), check::pair_str(#a, #b, (a), (b)), true)
#define EXPECT_LT(a, b) GC_CHECK_((a) < (b), check::pair_str(#a, #b, (a), (b)), false)
#define EXPECT_LE(a, b) GC_CHECK_((a) <= (b), check::pair_str(#a, #b, (a), (b)), false)
#define EXPECT_GT(a, b) GC_CHECK_((a) > (b), check::pair_str(#a, #b, (a), (b)), false)
#define EXPECT_NEAR(a, b, tol) \
  GC_CHECK_(std::abs((a) - (b)) <= (tol), check::pair_str(#a, #b, (a), (b)), false)
#define EXPECT_THROW_MSG(stmt, Exc, substr)                                             \
  do {                                                                                  \
    bool caught_ = false;                                                               \
    try {                                                                               \
      stmt;                                                                             \
    } catch (const Exc& e_) {                                                           \
      caught_ = std::string(e_.what()).find(substr) != std::string::npos;               \
      if (!caught_) check::fail(__FILE__, __LINE__, std::string("message: ") + e_.what(), false); \
      else caught_ = true;                                                              \
    }                                                                                   \
    if (!caught_) check::fail(__FILE__, __LINE__, "expected " #Exc " from " #stmt, false); \
  } while (0)

#define CHECK_MAIN()                                                          \
  int main() {                                                                \
    int ran = 0;                                                              \
    for (auto& c : check::registry()) {                                       \
      const int before = check::failures();                                   \
      std::printf("[ RUN  ] %s\n", c.name.c_str());                           \
      std::fflush(stdout);                                                    \
      try {                                                                   \
        c.fn();                                                               \
      } catch (const check::Abort&) {                                         \
      } catch (const std::exception& e) {                                     \
        check::fail(__FILE__, __LINE__, std::string("exception: ") + e.what(), false); \
      }                                                                       \
      std::printf("[ %s ] %s\n", check::failures() == before ? " OK " : "FAIL", c.name.c_str()); \
      std::fflush(stdout);                                                    \
      ++ran;                                                                  \
    }                                                                         \
    std::printf("%d cases, %d failures\n", ran, check::failures());           \
    return check::failures() == 0 ? 0 : 1;                                    \
  }
