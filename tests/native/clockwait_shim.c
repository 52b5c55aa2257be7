/* TEST INFRASTRUCTURE, sanitizer builds only. libstdc++ waits with a steady-clock timeout
   (condition_variable::wait_for / wait_until) through pthread_cond_clockwait, which this
   toolchain's ThreadSanitizer does not intercept: it then misses the wait's unlock and
   re-lock of the mutex and reports a "double lock" when another thread takes that mutex.
   The test programs built with the sanitizers define pthread_cond_clockwait themselves
   (exported, so libgail.so binds to it) as the same wait through pthread_cond_timedwait,
   which ThreadSanitizer does intercept, with the deadline moved to CLOCK_REALTIME. */
#define _GNU_SOURCE
#include <pthread.h>
#include <time.h>

int pthread_cond_clockwait(pthread_cond_t* cond, pthread_mutex_t* mutex, clockid_t clock, const struct timespec* abstime) {
    struct timespec now, real;
    clock_gettime(clock, &now);
    clock_gettime(CLOCK_REALTIME, &real);
    long long ns = (long long)(abstime->tv_sec - now.tv_sec) * 1000000000LL + (abstime->tv_nsec - now.tv_nsec);
    if (ns < 0) ns = 0;
    ns += real.tv_nsec;
    real.tv_sec += (time_t)(ns / 1000000000LL);
    real.tv_nsec = (long)(ns % 1000000000LL);
    return pthread_cond_timedwait(cond, mutex, &real);
}
