// Per-read seed lookup and reverse complement, shared by the gfx950 kernels and
// the host test build (same source, so the CPU suite checks the kernel logic).
#pragma once
#include <stdint.h>
#include "ga_device_types.h"

#if defined(__HIPCC__)
#define GA_SEED_HD __host__ __device__ inline
#else
#define GA_SEED_HD inline
#endif

namespace gad {

GA_SEED_HD int readBaseCode(unsigned char c)
{
	switch (c)
	{
		case 'a': case 'A': return 0;
		case 'c': case 'C': return 1;
		case 'g': case 'G': return 2;
		case 't': case 'T': return 3;
	}
	return -1;
}

// The seeder's 64-bit k-mer hash (src/MinimizerSeeder.cpp:45-54).
GA_SEED_HD uint64_t seedKmerHash(uint64_t key)
{
	key = (~key) + (key << 21);
	key = key ^ (key >> 24);
	key = (key + (key << 3)) + (key << 8);
	key = key ^ (key >> 14);
	key = (key + (key << 2)) + (key << 4);
	key = key ^ (key >> 28);
	key = key + (key << 31);
	return key;
}

// Whether the seeder's iterateKmers (src/MinimizerSeeder.cpp:60-102) reports the
// k-mer ending at read position i, and that k-mer. iterateKmers walks each run of
// ACGT letters (it restarts after any other letter) and reports a k-mer unless
// it equals the last reported one and lies less than a window of w - k + 1
// positions after it. Consecutive k-mers are equal only inside a run of one base
// longer than k, so the rule is a function of the position: k-mers are reported
// where they differ from the previous one, and inside such a run at its first
// k-mer and every w - k + 1 positions after it. Each position can thus be
// decided on its own, as the block-parallel kernel does.
GA_SEED_HD bool kmerReported(const char* s, uint64_t i, uint32_t k, uint32_t w, uint64_t& kmer)
{
	if (i + 1 < k) return false;
	kmer = 0;
	for (uint64_t p = i + 1 - k; p <= i; p++)
	{
		const int c = readBaseCode((unsigned char)s[p]);
		if (c < 0) return false;
		kmer = (kmer << 2) | (uint64_t)c;
	}
	// length of the run of one base ending at i
	const int c = readBaseCode((unsigned char)s[i]);
	uint64_t h = 1;
	while (h <= i && readBaseCode((unsigned char)s[i - h]) == c) h++;
	if (h <= k) return true;
	const uint64_t realWindow = w - k + 1;
	return (h - k) % realWindow == 0;
}

// Table slot of kmer (open addressing in seedKmerHash order), or null.
GA_SEED_HD const SeedTableSlot* seedTableFind(const SeedTableSlot* table, uint64_t mask, uint64_t kmer)
{
	uint64_t slot = seedKmerHash(kmer) & mask;
	while (true)
	{
		const uint64_t key = table[slot].kmer;
		if (key == kmer) return &table[slot];
		if (key == ~0ull) return nullptr;
		slot = (slot + 1) & mask;
	}
}

// Looks up every read k-mer that iterateKmers reports; k-mers present in the
// table with fewer than maxCount positions are written to out in read order.
// Returns their number.
GA_SEED_HD uint32_t lookupReadKmers(const char* s, uint64_t len, const SeedTableSlot* table, uint64_t mask, uint32_t k, uint32_t w, uint64_t maxCount, SeedMatch* out)
{
	uint32_t n = 0;
	for (uint64_t i = 0; i < len; i++)
	{
		uint64_t kmer;
		if (!kmerReported(s, i, k, w, kmer)) continue;
		const SeedTableSlot* t = seedTableFind(table, mask, kmer);
		if (!t || t->count >= maxCount) continue;
		SeedMatch m;
		m.pos = (uint32_t)i;
		m.start = t->start;
		m.count = t->count;
		out[n++] = m;
	}
	return n;
}

// CommonUtils::Complement (src/CommonUtils.cpp:67-97); letters outside the
// table are kept as they are (the host drops such reads before they are used).
GA_SEED_HD char complementByte(char c)
{
	switch (c)
	{
		case 'A': case 'a': return 'T';
		case 'C': case 'c': return 'G';
		case 'T': case 't': return 'A';
		case 'G': case 'g': return 'C';
		case 'N': case 'n': return 'N';
		case 'U': case 'u': return 'A';
		case 'R': case 'r': return 'Y';
		case 'Y': case 'y': return 'R';
		case 'K': case 'k': return 'M';
		case 'M': case 'm': return 'K';
		case 'S': case 's': return 'S';
		case 'W': case 'w': return 'W';
		case 'B': case 'b': return 'V';
		case 'V': case 'v': return 'B';
		case 'D': case 'd': return 'H';
		case 'H': case 'h': return 'D';
	}
	return c;
}

} // namespace gad
