// Image-section hand-over over mray_amd::TimelineSemaphore, both sides:
//   producer (the tracer, RenderImage::TransferToHost): Acquire(v), write the
//     staging buffer, Release(), v += 2, publish waitCounter = v - 1;
//   consumer (the run command): Acquire(waitCounter), read, Release().
// Each section writes a distinct pattern; the consumer checks it sees every
// section once, in order, and never a partly overwritten buffer.
#include "mray_tracer.hpp"

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

int main()
{
    using mray_amd::TimelineSemaphore;
    constexpr int kSections = 1000;
    constexpr size_t kN = 4096;
    TimelineSemaphore sem(0);
    std::vector<uint32_t> staging(kN, 0);
    std::deque<uint64_t> q;
    std::mutex m;
    std::condition_variable cv;
    int bad = 0, seen = 0;
    std::thread consumer([&]
    {
        for(int k = 0; k < kSections; k++)
        {
            uint64_t wait;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return !q.empty(); });
                wait = q.front();
                q.pop_front();
            }
            if(!sem.Acquire(wait)) { bad++; return; }
            for(size_t i = 0; i < kN; i++)
                if(staging[i] != uint32_t(k) * 7919u + uint32_t(i)) { bad++; break; }
            if(k % 97 == 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
            seen++;
            sem.Release();
        }
    });
    uint64_t v = 0;
    for(int k = 0; k < kSections; k++)
    {
        if(!sem.Acquire(v)) { bad++; break; }
        for(size_t i = 0; i < kN; i++) staging[i] = uint32_t(k) * 7919u + uint32_t(i);
        sem.Release();
        v += 2;
        std::lock_guard<std::mutex> lk(m);
        q.push_back(v - 1);
        cv.notify_all();
    }
    consumer.join();
    // an invalidated semaphore releases a waiter with false
    TimelineSemaphore s2(0);
    bool got = true;
    std::thread w([&] { got = s2.Acquire(5); });
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    s2.Invalidate();
    w.join();
    std::printf("sections %d seen %d bad %d invalidated-acquire %d\n", kSections, seen, bad, int(got));
    if(bad == 0 && seen == kSections && !got) std::printf("ok\n");
    return (bad == 0 && seen == kSections && !got) ? 0 : 1;
}
