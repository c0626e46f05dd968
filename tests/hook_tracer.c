/* A minimal tracer for the GPU trace test: defines the reference's three trace
 * hooks (include/common.h:9-11) and writes one line per completed task,
 * "name cpu start_us end_us", to the file named by PALIAL_HOOK_OUT.  Loaded
 * with LD_PRELOAD in front of the palial driver. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static FILE *out(void)
{
    static FILE *f = NULL;
    if (!f) {
        const char *p = getenv("PALIAL_HOOK_OUT");
        f = fopen(p ? p : "/dev/null", "w");
    }
    return f;
}

static char pending[64];
static int cpu_of_pending = -100;

void palial_set_task_name(const char *name)
{
    snprintf(pending, sizeof pending, "%s", name);
}

void palial_set_task_cpu_node(int cpu, int node, char *name)
{
    (void)node;
    if (strcmp(name, pending) == 0) cpu_of_pending = cpu;
}

void palial_get_task_time(struct timeval start, struct timeval end, char *name)
{
    long long s = (long long)start.tv_sec * 1000000 + start.tv_usec;
    long long e = (long long)end.tv_sec * 1000000 + end.tv_usec;
    fprintf(out(), "%s %d %lld %lld\n", name, strcmp(name, pending) == 0 ? cpu_of_pending : -100, s, e);
    fflush(out());
}
