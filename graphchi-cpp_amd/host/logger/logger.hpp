// Levelled logging with the reference's interface: logstream(LOG_x) << ... ;
// a LOG_FATAL line throws when it is terminated with std::endl
// (src/logger/logger.hpp:71-76, 215-217).
#pragma once
#include <iostream>
#include <sstream>
#include <string>

#define LOG_DEBUG 1
#define LOG_INFO 2
#define LOG_WARNING 3
#define LOG_ERROR 4
#define LOG_FATAL 5

namespace graphchi {

inline int &log_level() {
    static int lvl = LOG_INFO;
    return lvl;
}

class log_line {
  public:
    explicit log_line(int level) : level_(level) {}
    template <typename T>
    log_line &operator<<(const T &v) {
        buf_ << v;
        return *this;
    }
    log_line &operator<<(std::ostream &(*f)(std::ostream &)) {
        buf_ << f;
        if (f == static_cast<std::ostream &(*)(std::ostream &)>(std::endl)) flush();
        return *this;
    }
    ~log_line() {
        if (!buf_.str().empty()) emit();
    }

  private:
    void emit() {
        static const char *names[] = {"", "DEBUG", "INFO", "WARNING", "ERROR", "FATAL"};
        if (level_ >= log_level()) std::cerr << names[level_] << ": " << buf_.str();
        buf_.str("");
    }
    void flush() {
        emit();
        if (level_ == LOG_FATAL) throw "log fatal";
    }
    int level_;
    std::ostringstream buf_;
};

inline log_line logstream(int level) { return log_line(level); }

}  // namespace graphchi
