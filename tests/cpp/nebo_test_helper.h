// Small pass/fail helper in the spirit of the reference's test/TestHelper.h.
#pragma once
#include <cmath>
#include <cstdio>
#include <string>

struct TestHelper {
  bool ok = true;
  void operator()(bool cond, const std::string& label)
  {
    std::printf("%s : %s\n", cond ? "PASS" : "FAIL", label.c_str());
    ok = ok && cond;
  }
  int report() const
  {
    std::printf(ok ? "ALL TESTS PASSED\n" : "AT LEAST ONE TEST FAILED\n");
    return ok ? 0 : 1;
  }
};
