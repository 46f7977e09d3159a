// Minimal console logger with the severities of the reference's plog use
// (PLOG_FATAL/WARNING/INFO/DEBUG): "<time> LEVEL message" on stdout.
#pragma once

#include <chrono>
#include <cstdio>
#include <ctime>
#include <sstream>
#include <string>

namespace upclog
{
enum Level { FATAL = 1, ERROR_ = 2, WARNING = 3, INFO = 4, DEBUG = 5, VERBOSE = 6 };

inline int& maxLevel()
{
  static int level = VERBOSE;
  return level;
}

class Line
{
 public:
  explicit Line(Level lv) : lv_(lv) {}
  ~Line()
  {
    if (lv_ > maxLevel()) return;
    static const char* names[] = {"", "FATAL", "ERROR", "WARN ", "INFO ", "DEBUG", "VERB "};
    const auto now = std::chrono::system_clock::now();
    const std::time_t t = std::chrono::system_clock::to_time_t(now);
    const int ms = (int)(std::chrono::duration_cast<std::chrono::milliseconds>(now.time_since_epoch()).count() % 1000);
    char ts[32];
    std::tm tmv{};
    localtime_r(&t, &tmv);
    std::strftime(ts, sizeof(ts), "%Y-%m-%d %H:%M:%S", &tmv);
    std::fprintf(stdout, "%s.%03d %s %s\n", ts, ms, names[lv_], os_.str().c_str());
    std::fflush(stdout);
  }
  template <class T>
  Line& operator<<(const T& v)
  {
    os_ << v;
    return *this;
  }

 private:
  Level lv_;
  std::ostringstream os_;
};
}  // namespace upclog

#define UPC_LOG_FATAL upclog::Line(upclog::FATAL)
#define UPC_LOG_ERROR upclog::Line(upclog::ERROR_)
#define UPC_LOG_WARNING upclog::Line(upclog::WARNING)
#define UPC_LOG_INFO upclog::Line(upclog::INFO)
#define UPC_LOG_DEBUG upclog::Line(upclog::DEBUG)
