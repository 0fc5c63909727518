// mini_test.hpp -- a few gtest-style macros (gtest is not installed here).
// Tests register with TEST(Suite, Name); main() runs those whose
// "Suite.Name" contains the filter argument and reports PASS/FAIL lines.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mini {

struct Failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Case {
    std::string name;
    std::function<void()> fn;
};

inline std::vector<Case>& registry()
{
    static std::vector<Case> r;
    return r;
}

struct Registrar {
    Registrar(const char* suite, const char* name, std::function<void()> fn)
    {
        registry().push_back({std::string(suite) + "." + name, std::move(fn)});
    }
};

// Equal within 4 ULPs in float, as gtest's ASSERT_FLOAT_EQ.
inline bool float_eq(float a, float b)
{
    if (std::isnan(a) || std::isnan(b)) return false;
    int32_t ia, ib;
    std::memcpy(&ia, &a, 4);
    std::memcpy(&ib, &b, 4);
    if (ia < 0) ia = INT32_MIN - ia;
    if (ib < 0) ib = INT32_MIN - ib;
    int64_t d = static_cast<int64_t>(ia) - ib;
    return (d < 0 ? -d : d) <= 4;
}

template <class A, class B> std::string show(const A& a, const B& b)
{
    std::ostringstream o;
    o.precision(17);
    o << a << " vs " << b;
    return o.str();
}

inline int run_all(int argc, char** argv)
{
    const char* filter = argc > 1 ? argv[1] : "";
    int failed = 0, ran = 0;
    for (auto& c : registry()) {
        if (c.name.find(filter) == std::string::npos) continue;
        ++ran;
        try {
            c.fn();
            std::printf("PASS %s\n", c.name.c_str());
        } catch (const std::exception& e) {
            ++failed;
            std::printf("FAIL %s: %s\n", c.name.c_str(), e.what());
        }
        std::fflush(stdout);
    }
    std::printf("%d run, %d failed\n", ran, failed);
    return failed ? 1 : (ran ? 0 : 2);
}

} // namespace mini

#define MINI_CAT2(a, b) a##b
#define MINI_CAT(a, b) MINI_CAT2(a, b)
#define TEST(suite, name)                                                              \
    static void suite##_##name();                                                      \
    static mini::Registrar MINI_CAT(reg_, MINI_CAT(suite, name))(#suite, #name,        \
                                                                 suite##_##name);      \
    static void suite##_##name()

#define MINI_FAIL(msg)                                                                 \
    throw mini::Failure(std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " + \
                        (msg))
#define ASSERT_TRUE(c)                                                                 \
    do {                                                                               \
        if (!(c)) MINI_FAIL("expected true: " #c);                                     \
    } while (0)
#define EXPECT_TRUE(c) ASSERT_TRUE(c)
#define ASSERT_EQ(a, b)                                                                \
    do {                                                                               \
        if (!((a) == (b))) MINI_FAIL(#a " == " #b ": " + mini::show((a), (b)));        \
    } while (0)
#define EXPECT_EQ(a, b) ASSERT_EQ(a, b)
#define ASSERT_LE(a, b)                                                                \
    do {                                                                               \
        if (!((a) <= (b))) MINI_FAIL(#a " <= " #b ": " + mini::show((a), (b)));        \
    } while (0)
#define ASSERT_GE(a, b)                                                                \
    do {                                                                               \
        if (!((a) >= (b))) MINI_FAIL(#a " >= " #b ": " + mini::show((a), (b)));        \
    } while (0)
#define ASSERT_FLOAT_EQ(a, b)                                                          \
    do {                                                                               \
        if (!mini::float_eq(static_cast<float>(a), static_cast<float>(b)))             \
            MINI_FAIL(#a " ~= " #b ": " + mini::show((a), (b)));                       \
    } while (0)
#define ASSERT_NEAR(a, b, tol)                                                         \
    do {                                                                               \
        if (!(std::fabs((a) - (b)) <= (tol)))                                          \
            MINI_FAIL(#a " near " #b ": " + mini::show((a), (b)));                     \
    } while (0)
#define EXPECT_THROW_MSG(stmt, exc, text)                                              \
    do {                                                                               \
        bool thrown_ = false;                                                          \
        try {                                                                          \
            stmt;                                                                      \
        } catch (const exc& e_) {                                                      \
            thrown_ = true;                                                            \
            if (std::string(e_.what()) != (text))                                      \
                MINI_FAIL(std::string("message: ") + e_.what());                       \
        }                                                                              \
        if (!thrown_) MINI_FAIL("expected " #exc " from " #stmt);                      \
    } while (0)
