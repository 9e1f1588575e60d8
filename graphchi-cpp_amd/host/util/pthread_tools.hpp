// Lock types of the program API (src/util/pthread_tools.hpp:36-157): programs
// such as trianglecounting keep host-side state behind a graphchi::mutex.
#pragma once
#include <pthread.h>

#include <cstdio>
#include <cstdlib>

namespace graphchi {

class mutex {
  public:
    mutex() {
        if (pthread_mutex_init(&m_, nullptr) != 0) {
            std::perror("graphchi::mutex: pthread_mutex_init");
            std::abort();
        }
    }
    ~mutex() { pthread_mutex_destroy(&m_); }
    mutex(const mutex &) = delete;
    mutex &operator=(const mutex &) = delete;
    void lock() const { pthread_mutex_lock(&m_); }
    void unlock() const { pthread_mutex_unlock(&m_); }
    bool try_lock() const { return pthread_mutex_trylock(&m_) == 0; }

  private:
    mutable pthread_mutex_t m_;
};

// Same interface; the host layer runs programs' callbacks on one thread, so a
// plain mutex serves.
typedef mutex spinlock;

}  // namespace graphchi
