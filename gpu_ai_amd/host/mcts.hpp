// mcts.hpp -- the batch UCB1 search tree of gpu_ai (src/mcts.hpp:12-64), restated.  select()
// hands out a whole batch of trials at once and returns one State per trial; update() takes the
// playout winners in the same order.  The allocation rule, UCB1 arithmetic and scoring follow
// src/mcts.cpp:63-191, so a tree grown with the same results makes the same choices; win counts
// are doubles so that leaves scored by a value network can add fractional results.
#pragma once

#include <memory>
#include <vector>

#include "playout.hpp"
#include "state.hpp"

class GameTree {
 public:
  explicit GameTree(State state, GameTree *parent = nullptr);
  ~GameTree();

  // Performs a move, keeping the matching subtree (src/mcts.cpp:11-25).
  std::shared_ptr<GameTree> move(const Move &move);
  // Fraction of wins for a player; a finished game scores 1 for the side not to move
  // (src/mcts.cpp:27-37).
  double getScore(PlayerId player) const;
  // Child move of highest score for a player (src/mcts.cpp:39-55).
  Move getOptMove(PlayerId player) const;
  unsigned getTotalTrials() const { return totalTrials; }
  // select + runPlayouts + update (src/mcts.cpp:57-61).
  void expand(unsigned numPlayouts, PlayoutDriver *playoutDriver);
  // Distributes `trials` by UCB1, creating nodes as needed (src/mcts.cpp:63-157).
  std::vector<State> select(unsigned trials);
  // As select, appending the trial states to `out` (no per-node allocation).
  void selectInto(unsigned trials, std::vector<State> &out);
  // Adds the winners of the trials handed out by the last select (src/mcts.cpp:159-180).
  void update(const std::vector<PlayerId> &results);
  void update(const PlayerId *results, size_t n);
  // As update, with each trial scored by player 1's win probability p: player 1 gains p and
  // player 2 gains 1 - p (value-network leaves).
  void updateValues(const std::vector<float> &p1WinProb);
  void updateValues(const float *p1WinProb, size_t n);
  // Whether select parallelises over the root's children with OpenMP (as the reference does at
  // every node, src/mcts.cpp:146-149); searches run in parallel with each other switch it off.
  void setParallel(bool on) { parallel = on; }
  // The factor k of the exploration term sqrt(k ln N / n); the reference's UCB1 has k = 2
  // (src/mcts.cpp:189-190), the default.  Nodes created later inherit it.
  void setExploration(double k) { explore = k; }
  // Whose win fraction the exploitation term counts.  The reference's UCB1 counts the wins of the
  // side to move in the child (src/mcts.cpp:189), that is, of the opponent of the player who
  // chooses among the children, so its trials lean towards moves that are good for the opponent;
  // on, the term counts the chooser's own wins.  Off by default (the reference's arithmetic);
  // nodes created later inherit it.
  void setMoverWins(bool on) { moverWins = on; }

  const State state;

  // introspection
  const std::vector<std::shared_ptr<GameTree>> &getChildren() const { return children; }
  const std::vector<Move> &getMoves() const;  // generated on first use
  double getWins(PlayerId p) const { return wins[p]; }
  unsigned getAssignedTrials() const { return assignedTrials; }
  bool isExpanded() const { return expanded; }

 private:
  double ucb1() const;  // src/mcts.cpp:182-191
  double ucb1(double lnParent, double k, PlayerId side) const;
  template <class R, class Add>
  void updateImpl(const R *results, size_t n, Add add, double sums[NUM_PLAYERS]);

  const GameTree *parent;
  bool expanded = false;
  bool over = false;       // isGameOver() of the node's state, fixed at construction
  bool parallel = false;
  double explore = 2.0;
  bool moverWins = false;
  mutable std::vector<Move> moves;
  mutable bool movesReady = false;
  std::vector<std::shared_ptr<GameTree>> children;
  unsigned assignedTrials = 0;
  unsigned totalTrials = 0;
  double wins[NUM_PLAYERS] = {0.0, 0.0};
};
